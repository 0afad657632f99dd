// One-time correlators and the single-particle Green's function on the device
// (c++/triqs_xca/block_sparse_backbone.cpp:526-706,732-902; python/triqs_xca/block_sparse_solver.py:657-698):
// plan construction, launches and the C ABI.
//
// Per (direction pattern, orbital tuple, rule class) group, the chain kernel runs the [tau, 0] side
// of every used block path and the transposed [beta, tau] side in store mode, writing one
// (r, rows, cols) array per pole tuple; refl_kernel reflects the [beta, tau] outputs in place,
// corr_kernel contracts matching pairs with O_mu and O_kap per tuple and time node, and
// corr_reduce_kernel sums the chunk slots in a fixed order. Groups are cut into batches whose stored
// outputs fit a bounded buffer.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/xca_amd.h"
#include "ctx.hpp"
#include "engine.hpp"
#include "host/errors.hpp"
#include "host/model.hpp"
#include "plan.hpp"

using namespace xca;

namespace {

int64_t ipow(int64_t a, int e) {
  int64_t r = 1;
  for (int i = 0; i < e; ++i) r *= a;
  return r;
}

// c_o and c^dag_o per orbital, in orbital order (setup_mu_ops / setup_kap_ops_for_single_ptcle_gf,
// block_sparse_backbone.cpp:737-777)
CorrOps spgf_ops(const Model &M) {
  const Problem &P = M.P;
  const int n = P.n, nb = P.nb;
  CorrOps co;
  co.nmu = co.nkap = n;
  co.mu_conn = P.c_conn;
  co.kap_conn = P.cdag_conn;
  co.mu_off.assign((size_t)n * nb, -1);
  co.kap_off.assign((size_t)n * nb, -1);
  for (int dag = 0; dag < 2; ++dag)
    for (int o = 0; o < n; ++o)
      for (int j = 0; j < nb; ++j) {
        const auto &conn = dag ? P.cdag_conn : P.c_conn;
        if (conn[(size_t)o * nb + j] < 0) continue;
        const auto &mat = (dag ? P.cdag_mats : P.c_mats)[(size_t)o * nb + j];
        (dag ? co.kap_off : co.mu_off)[(size_t)o * nb + j] = (int64_t)co.mats.size();
        co.mats.insert(co.mats.end(), mat.begin(), mat.end());
      }
  return co;
}

// operators given as (nops, nblocks) connections and matrices packed in (op, source block) order
void add_ops(const Problem &P, int nops, const int *conn, const double *mats, std::vector<int> &c_out, std::vector<int64_t> &off,
             std::vector<double> &all, const char *what) {
  if (nops <= 0) throw std::invalid_argument(std::string("xca_correlator: no ") + what + " operators");
  if (!conn || !mats) throw std::invalid_argument(std::string("xca_correlator: null ") + what + " operator arrays");
  const int nb = P.nb;
  c_out.assign(conn, conn + (size_t)nops * nb);
  off.assign((size_t)nops * nb, -1);
  int64_t pos = 0;
  for (int o = 0; o < nops; ++o)
    for (int j = 0; j < nb; ++j) {
      const int t = c_out[(size_t)o * nb + j];
      if (t < -1 || t >= nb) throw std::invalid_argument(std::string("xca_correlator: ") + what + " operator block index out of range");
      if (t < 0) continue;
      const int64_t sz = (int64_t)P.dims[t] * P.dims[j];
      off[(size_t)o * nb + j] = (int64_t)all.size();
      for (int64_t x = 0; x < sz; ++x) {
        const double v = mats[pos + x];
        if (!std::isfinite(v)) throw std::invalid_argument(std::string("xca_correlator: non-finite ") + what + " operator entry");
        all.push_back(v);
      }
      pos += sz;
    }
}

// doubles of stored per-item outputs one batch may hold (XCA_CORR_STORE_MB overrides)
int64_t store_budget() {
  if (const char *e = std::getenv("XCA_CORR_STORE_MB")) {
    const long v = std::atol(e);
    if (v > 0) return (int64_t)v * (1 << 20) / 8;
  }
  return (int64_t)1 << 29;  // 4 GiB
}

// one host group with its tuple selection: first-line poles [a, b) of the class enumeration, or the
// single tuple u (explicit flat-index subsets)
struct DevGroup {
  int g = 0, a = 0, b = 0, u = -1;
  int64_t ntup = 0;
};

}  // namespace

namespace xca {

void corr_release(xca_ctx *c) { c->corr_plans.clear(); }

static CorrDevPlan *get_corr_plan(xca_ctx *c, int m, const int *topo, const int32_t *fix, int64_t nfix, const CorrOps &co, bool spgf) {
  const Problem &P = c->M.P;
  const int r = P.r, p = P.p, n = P.n, nb = P.nb;
  std::vector<int> key;
  key.push_back(m);
  key.insert(key.end(), topo, topo + 2 * m);
  key.insert(key.end(), c->G_zero.begin(), c->G_zero.end());
  key.push_back(spgf ? -2 : -3);
  if (!spgf) {  // connections and the nonzero pattern of the operators
    key.push_back(co.nmu);
    key.push_back(co.nkap);
    key.insert(key.end(), co.mu_conn.begin(), co.mu_conn.end());
    key.insert(key.end(), co.kap_conn.begin(), co.kap_conn.end());
    for (size_t x = 0; x < co.mats.size(); x += 32) {
      int bits = 0;
      for (size_t y = x; y < std::min(co.mats.size(), x + 32); ++y) bits |= (co.mats[y] != 0.0 ? 1 : 0) << (y - x);
      key.push_back(bits);
    }
  }
  if (fix) {
    key.push_back(-1);
    key.insert(key.end(), fix, fix + nfix);
  }
  auto found = c->corr_plans.find(key);
  if (found != c->corr_plans.end()) return found->second.get();
  if (fix)  // one explicit-subset plan at a time
    for (auto e = c->corr_plans.begin(); e != c->corr_plans.end();) {
      if (std::find(e->first.begin(), e->first.end(), -1) == e->first.end()) {
        ++e;
        continue;
      }
      HIP_CHECK(hipStreamSynchronize(c->stream));  // its launches may still be in flight
      e = c->corr_plans.erase(e);
    }

  if (m < 1 || m > kMaxLines + 1) throw std::invalid_argument("xca: diagram order must be between 1 and 4");
  CorrPlanHost hp;
  build_corr_programs(c->M, c->ps, topo, m, c->G_zero, co, hp);
  int variant = c->tile_variant >= 0 ? c->tile_variant : 0;
  int max_w = 0;
  // a context with blocks above the tile engine's (or XCA_WIDE=1) runs the chains on the large-block
  // engine, as its Σ evaluations do
  const bool wide = c->wide;
  for (auto &pg : hp.progs) {
    if (!wide)
      for (auto &op : pg.ops)
        if (op.bin > 12 || op.bout > 12) throw std::logic_error("xca: block above 12 rows in a tile-engine context");
    max_w = std::max(max_w, pg.max_w);
  }
  if (!wide && max_w > tile_columns(variant)) variant = 0;
  if (!wide && max_w > tile_columns(variant)) throw std::invalid_argument("xca: block path too wide for one tile");
  const int NT = tile_columns(variant);
  auto dp = std::make_unique<CorrDevPlan>();
  dp->nmu = co.nmu;
  dp->nkap = co.nkap;
  dp->variant = variant;
  const int nl = m - 1;
  const int nout = co.nmu * co.nkap;

  // device ops and per-program templates
  std::vector<DOp> h_ops;
  std::vector<DProg> base(hp.progs.size());
  std::vector<int64_t> per_item(hp.progs.size());  // stored doubles per pole tuple
  for (size_t i = 0; i < hp.progs.size(); ++i) {
    const Program &pg = hp.progs[i];
    DProg d{};
    d.op0 = (int)h_ops.size();
    d.nops = (int)pg.ops.size();
    if (!wide)
      for (auto &op : pg.ops) h_ops.push_back(device_op(op, hp.tables, c));
    d.sigma_block = -1;
    d.b0 = pg.out_cols;
    d.b1 = pg.b1;
    d.g_init = c->M.G_off[pg.init_block];
    d.sign = pg.sign;
    d.nlines = nl;
    d.max_w = pg.max_w;
    d.split = 0;
    d.share_len = 1;
    d.gtr = pg.gtr;
    int fb[kMaxLines + 1];
    for (int i2 = 1, f = pg.fb_ix; i2 < m; ++i2, f /= 2) fb[i2] = f % 2;
    for (int x = 0; x < nl; ++x) {
      d.pk_sign[x] = pg.pk_sign[x];
      d.pk_exp[x] = pg.pk_exp[x];
      const int fwd = fb[x + 1], rb = (pg.cls >> x) & 1;
      d.set_off[x] = c->plist_off[(fwd ? 0 : 2) + rb];
      d.set_len[x] = (int)c->ps.get(fwd, rb).size();
    }
    d.items_per_tile = std::max(1, NT / d.max_w);
    base[i] = d;
    per_item[i] = (int64_t)r * pg.out_rows * pg.out_cols;
  }
  auto class_tuples = [&](const DProg &d) {
    int64_t t = 1;
    for (int x = 0; x < nl; ++x) t *= d.set_len[x];
    return t;
  };

  // device groups
  const int64_t budget = store_budget();
  std::vector<DevGroup> dgroups;
  if (!fix) {
    for (size_t gi = 0; gi < hp.groups.size(); ++gi) {
      const CorrGroup &g = hp.groups[gi];
      const DProg &d0 = base[g.prog0];
      const int64_t ntup = class_tuples(d0);
      int64_t st = 0;
      for (int pi = g.prog0; pi < g.prog1; ++pi) st += per_item[pi];
      const int L0 = nl > 0 ? d0.set_len[0] : 1;
      const int64_t inner = ntup / L0;  // tuples per first-line pole
      const int step = (int)std::max<int64_t>(1, std::min<int64_t>(L0, budget / std::max<int64_t>(1, st * inner)));
      for (int a = 0; a < L0; a += step) {
        DevGroup dg;
        dg.g = (int)gi;
        dg.a = a;
        dg.b = std::min(L0, a + step);
        dg.ntup = (int64_t)(dg.b - dg.a) * inner;
        dgroups.push_back(dg);
      }
    }
  } else {
    const int64_t o_max = ipow(n, m - 1), p_max = ipow(p, m - 1), f_max = ipow(2, m - 1) * o_max * p_max;
    std::map<std::pair<int, int>, std::vector<int>> wanted;  // (fb_ix, o_ix) -> p_ix list
    for (int64_t i = 0; i < nfix; ++i) {
      const int64_t f = fix[i];
      // the reference decodes any integer; an index past the end would alias another backbone
      if (f < 0 || f >= f_max) throw std::runtime_error("DiagramEvaluator::eval_correlator: f_ix out of range");
      const int o_ix = (int)(f % o_max);
      const int64_t rest = f / o_max;
      wanted[{(int)(rest / p_max), o_ix}].push_back((int)(rest % p_max));
    }
    for (size_t gi = 0; gi < hp.groups.size(); ++gi) {
      const CorrGroup &g = hp.groups[gi];
      auto w = wanted.find({g.fb_ix, g.o_ix});
      if (w == wanted.end()) continue;
      int fb[kMaxLines + 1];
      for (int i2 = 1, f = g.fb_ix; i2 < m; ++i2, f /= 2) fb[i2] = f % 2;
      for (int pix : w->second) {
        int u = 0, stride = 1, cls = 0, pp = pix;
        for (int x = 0; x < nl; ++x) {
          const int l = pp % p;
          pp /= p;
          const int fwd = fb[x + 1];
          const double wl = P.poles[l] * P.beta;
          const int rb = fwd ? (wl <= 0 ? 0 : 1) : (wl >= 0 ? 0 : 1);  // backbone.cpp:88,110
          cls |= rb << x;
          const auto &lst = c->ps.get(fwd, rb);
          u += (int)(std::find(lst.begin(), lst.end(), l) - lst.begin()) * stride;
          stride *= (int)lst.size();
        }
        if (cls != g.cls) continue;
        DevGroup dg;
        dg.g = (int)gi;
        dg.u = u;
        dg.ntup = 1;
        dgroups.push_back(dg);
      }
    }
  }

  // operator blocks as sparse rows: cidx holds row pointers then column indices, cval_src the
  // position in co.mats of every stored value
  std::map<int64_t, std::pair<int, int>> csr_of;  // dense offset -> (cidx offset, value offset)
  std::vector<int> h_cidx;
  auto csr = [&](int64_t dense, int rows, int cols) {
    auto f = csr_of.find(dense);
    if (f != csr_of.end()) return f->second;
    std::pair<int, int> o{(int)h_cidx.size(), (int)dp->cval_src.size()};
    std::vector<int> colv;
    h_cidx.push_back(0);
    for (int i = 0; i < rows; ++i) {
      for (int j = 0; j < cols; ++j)
        if (co.mats[dense + (int64_t)i * cols + j] != 0.0) {
          colv.push_back(j);
          dp->cval_src.push_back(dense + (int64_t)i * cols + j);
        }
      h_cidx.push_back((int)colv.size());
    }
    h_cidx.insert(h_cidx.end(), colv.begin(), colv.end());
    csr_of[dense] = o;
    return o;
  };

  // batches
  int64_t max_store = 0, max_partial = 0;
  std::vector<WideCorrProg> wide_progs;
  size_t gi0 = 0;
  while (gi0 < dgroups.size()) {
    // groups [gi0, gi1) fill one batch
    size_t gi1 = gi0;
    int64_t store = 0;
    while (gi1 < dgroups.size()) {
      const DevGroup &dg = dgroups[gi1];
      const CorrGroup &g = hp.groups[dg.g];
      int64_t gs = 0;
      for (int pi = g.prog0; pi < g.prog1; ++pi) gs += per_item[pi] * dg.ntup;
      if (gi1 > gi0 && store + gs > budget) break;
      store += gs;
      ++gi1;
    }
    auto B = std::make_unique<CorrBatchDev>();
    std::vector<CTerm> terms;
    std::vector<CRun> runs;
    std::vector<CChunk> cchunks;
    std::vector<RJob> rjobs;
    int64_t off = 0, batch_items = 0;
    for (size_t k = gi0; k < gi1; ++k) batch_items += dgroups[k].ntup;
    const int kc = (int)std::max<int64_t>(1, std::min<int64_t>(64, batch_items / 4096));
    int64_t est_tiles = 0;
    for (size_t k = gi0; k < gi1; ++k) {
      const DevGroup &dg = dgroups[k];
      const CorrGroup &g = hp.groups[dg.g];
      std::map<int, int64_t> out_of;  // host program -> store offset
      for (int pi = g.prog0; pi < g.prog1; ++pi) {
        DProg d = base[pi];
        if (dg.u >= 0) {
          int decode = dg.u;
          for (int x = 0; x < nl; ++x) {
            const int len = d.set_len[x];
            d.set_off[x] += decode % len;
            decode /= len;
            d.set_len[x] = 1;
          }
        } else if (nl > 0) {
          d.set_off[0] += dg.a;
          d.set_len[0] = dg.b - dg.a;
        }
        d.ntup = (int)dg.ntup;
        d.out_off = off;
        out_of[pi] = off;
        if (wide) {  // the items in the chain kernel's store order: last line fastest
          WideCorrProg w{&hp.progs[pi], {}, off, (int)dp->batches.size()};
          w.poles.reserve((size_t)dg.ntup);
          for (int64_t u = 0; u < dg.ntup; ++u) {
            std::vector<int> ls(nl);
            int64_t v = u;
            for (int x = nl - 1; x >= 0; --x) {
              ls[x] = c->plist_h[d.set_off[x] + (int)(v % d.set_len[x])];
              v /= d.set_len[x];
            }
            w.poles.push_back(std::move(ls));
          }
          wide_progs.push_back(std::move(w));
        }
        if (hp.progs[pi].refl) {  // reflection jobs over the program's stored items
          const int w = hp.progs[pi].out_rows * hp.progs[pi].out_cols;
          const int64_t ncols = dg.ntup * w;
          for (int64_t c0 = 0; c0 < ncols; c0 += refl_columns())
            rjobs.push_back(RJob{off, w, (int)c0, (int)std::min<int64_t>(refl_columns(), ncols - c0), 0});
        }
        off += per_item[pi] * dg.ntup;
        if (!wide) B->h_progs.push_back(d);
        est_tiles += (d.ntup + d.items_per_tile - 1) / d.items_per_tile;
        dp->items += dg.ntup;
        dp->macs += canonical_macs(hp.progs[pi], r) * (double)dg.ntup;
      }
      const int run0 = (int)runs.size();
      for (size_t t = 0; t < g.terms.size(); ++t) {
        const CorrTerm &u = g.terms[t];
        CTerm ct{};
        ct.r_off = out_of[u.rprog];
        ct.l_off = out_of[u.lprog];
        ct.d0 = P.dims[u.e0];
        ct.d1 = P.dims[u.e1];
        ct.db = P.dims[u.b];
        ct.dc = P.dims[u.c];
        const int64_t mo = co.mu_off[(size_t)u.mu * nb + u.e0], ko = co.kap_off[(size_t)u.kap * nb + u.c];
        if (mo < 0 || ko < 0) throw std::logic_error("missing correlator operator block");
        const auto cm = csr(mo, ct.d1, ct.d0), ck = csr(ko, ct.db, ct.dc);
        ct.mu_csr = cm.first;
        ct.mu_val = cm.second;
        ct.kap_csr = ck.first;
        ct.kap_val = ck.second;
        const int out = u.mu * co.nkap + u.kap;
        if (runs.size() > (size_t)run0 && runs.back().out == out) runs.back().t1 = (int)terms.size() + 1;
        else runs.push_back(CRun{out, (int)terms.size(), (int)terms.size() + 1, 0});
        terms.push_back(ct);
        dp->macs += (double)r * ct.d1 * ct.db * (ct.d0 + ct.dc + 1) * (double)dg.ntup;
      }
      const int run1 = (int)runs.size();
      for (int64_t k0 = 0; k0 < dg.ntup; k0 += kc) cchunks.push_back(CChunk{run0, run1, (int)k0, (int)std::min<int64_t>(dg.ntup, k0 + kc)});
    }
    const int ct = chunk_tiles(est_tiles);
    for (size_t pi = 0; pi < B->h_progs.size(); ++pi) {
      const DProg &d = B->h_progs[pi];
      const int ntiles = (d.ntup + d.items_per_tile - 1) / d.items_per_tile;
      for (int t0 = 0; t0 < ntiles; t0 += ct) B->h_chunks.push_back(DChunk{(int)pi, t0, std::min(ntiles, t0 + ct), 0, 0});
    }
    costliest_first(B->h_chunks, B->h_progs);  // store mode: each item's output has its own place
    B->nchunks = (int)B->h_chunks.size();
    B->ncchunks = (int)cchunks.size();
    B->nrjobs = (int)rjobs.size();
    B->store_size = off;
    hipStream_t s = c->stream;
    B->progs.upload(B->h_progs, s);
    B->chunks.upload(B->h_chunks, s);
    B->terms.upload(terms, s);
    B->runs.upload(runs, s);
    B->cchunks.upload(cchunks, s);
    B->rjobs.upload(rjobs, s);
    max_store = std::max(max_store, off);
    max_partial = std::max(max_partial, (int64_t)B->ncchunks * r * nout);
    dp->device_bytes += (int64_t)(B->progs.n * sizeof(DProg) + B->chunks.n * sizeof(DChunk) + B->terms.n * sizeof(CTerm) +
                                  B->runs.n * sizeof(CRun) + B->cchunks.n * sizeof(CChunk) + B->rjobs.n * sizeof(RJob));
    dp->batches.push_back(std::move(B));
    gi0 = gi1;
  }
  hipStream_t s = c->stream;
  dp->ops.upload(h_ops, s);
  if (!wide) {
    dp->h_tab = tab_jobs(hp.tables, c);
    dp->tabjobs.upload(dp->h_tab, s);
    dp->ntabjobs = (int)dp->h_tab.size();
    dp->ftab.alloc(std::max<int64_t>(hp.ftab_size, 1));
  }
  dp->store.alloc(std::max<int64_t>(max_store, 1));
  dp->partial.alloc(std::max<int64_t>(max_partial, 1));
  dp->cidx.upload(h_cidx, s);
  dp->cval.alloc(std::max<size_t>(dp->cval_src.size(), 1));
  dp->queue.alloc(2);
  HIP_CHECK(init_queue(dp->queue.ptr, s));
  dp->device_bytes += (int64_t)(dp->ops.n * sizeof(DOp) + dp->tabjobs.n * sizeof(DTabJob)) +
                      8 * (int64_t)(dp->ftab.n + dp->store.n + dp->partial.n + dp->cval.n) + 4 * (int64_t)dp->cidx.n;
  if (wide) {
    dp->wide = build_wide_corr(c, hp.tables, hp.ftab_size, wide_progs, (int)dp->batches.size(), dp->store.ptr, (int64_t)dp->store.n);
    dp->device_bytes += dp->wide->device_bytes;
    dp->variant = -1;
  }
  // launch descriptors
  for (auto &B : dp->batches) {
    ChainArgs A{};
    A.D = c->D;
    A.ops = dp->ops.ptr; A.progs = B->progs.ptr; A.chunks = B->chunks.ptr;
    A.plist = c->plist.ptr; A.G = c->G.ptr; A.opmats = c->opmats.ptr; A.ftab = dp->ftab.ptr; A.kp = c->kp.ptr; A.km = c->km.ptr;
    A.kpow = c->kpow.ptr; A.partial = nullptr; A.store = dp->store.ptr;
    A.nchunks = B->nchunks;
    A.r = r; A.p = p; A.maxexp = c->maxexp;
    B->seg.upload(std::vector<ChainArgs>(1, A), s);
    B->seg_prefix.upload(std::vector<int>{0, B->nchunks}, s);
  }
  if (dp->ntabjobs > 0) {
    TabSeg T{};
    T.D = c->D;
    T.jobs = dp->tabjobs.ptr;
    T.G = c->G.ptr;
    T.kp = c->kp.ptr;
    T.km = c->km.ptr;
    T.ftab = dp->ftab.ptr;
    T.p = p;
    dp->tseg.upload(std::vector<TabSeg>(1, T), s);
    dp->tseg_prefix.upload(std::vector<int>{0, dp->ntabjobs}, s);
  }
  HIP_CHECK(hipStreamSynchronize(s));
  CorrDevPlan *raw = dp.get();
  c->corr_plans[key] = std::move(dp);
  return raw;
}

// Correlator of one topology into out (r, nmu, nkap); the chain-kernel time goes to the context's
// statistics.
static void corr_eval(xca_ctx *c, const double *G, int m, const int *topo, const int32_t *fix, int64_t nfix, const CorrOps &co, bool spgf,
                      double *out) {
  if (!c) throw std::invalid_argument("xca: null context");
  if (!topo || !out) throw std::invalid_argument("xca_correlator: null argument");
  if (fix && nfix < 0) throw std::invalid_argument("xca_correlator: negative f_ix count");
  if (G) {
    if (xca_set_G(c, G) < 0) throw std::runtime_error(xca_last_error());
  } else if (!c->have_G) {
    throw std::runtime_error("xca: G has not been set");
  }
  CorrDevPlan *dp = get_corr_plan(c, m, topo, fix, nfix, co, spgf);
  const Problem &P = c->M.P;
  const int r = P.r, nout = dp->nmu * dp->nkap;
  hipStream_t s = c->stream;
  c->corr_out.alloc((size_t)r * nout);
  std::vector<double> cval(dp->cval_src.size());
  for (size_t x = 0; x < cval.size(); ++x) cval[x] = co.mats[dp->cval_src[x]];
  if (!cval.empty()) HIP_CHECK(hipMemcpyAsync(dp->cval.ptr, cval.data(), cval.size() * sizeof(double), hipMemcpyHostToDevice, s));
  if (dp->ntabjobs > 0) HIP_CHECK(launch_tab(c->RB, TabBatch{dp->tseg.ptr, dp->tseg_prefix.ptr, 1, dp->ntabjobs}, s));
  if (dp->wide) wide_launch_range(c, *dp->wide, 0, dp->wide->seg[0]);
  std::vector<hipEvent_t> ev;
  for (size_t bi = 0; bi < dp->batches.size(); ++bi) {
    CorrBatchDev &B = *dp->batches[bi];
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    ev.push_back(e0);
    HIP_CHECK(hipEventCreate(&e1));
    ev.push_back(e1);
    HIP_CHECK(hipEventRecord(e0, s));
    if (dp->wide) {
      wide_launch_range(c, *dp->wide, dp->wide->seg[bi], dp->wide->seg[bi + 1]);
    } else {
      ChainBatch CB{B.seg.ptr, B.seg_prefix.ptr, 1, B.nchunks, dp->queue.ptr, nullptr, 0};
      CB.maxb = ctx_max_block(c);
      HIP_CHECK(launch_chain(c->RB, dp->variant, CB, s));
    }
    HIP_CHECK(hipEventRecord(e1, s));
    HIP_CHECK(launch_refl(c->RB, ReflArgs{c->D, B.rjobs.ptr, dp->store.ptr, B.nrjobs, 0}, s));
    CorrArgs A{B.terms.ptr, B.runs.ptr, B.cchunks.ptr, dp->store.ptr, dp->cval.ptr, dp->cidx.ptr, dp->partial.ptr, B.ncchunks, r, nout,
               dp->wide ? 1 : 0};
    HIP_CHECK(launch_corr(A, s));
    HIP_CHECK(launch_corr_reduce(dp->partial.ptr, B.ncchunks, r * nout, c->corr_out.ptr, bi > 0 ? 1 : 0, s));
  }
  if (dp->batches.empty()) HIP_CHECK(launch_fill((int64_t)r * nout, 0.0, c->corr_out.ptr, s));
  HIP_CHECK(hipMemcpyAsync(out, c->corr_out.ptr, (size_t)r * nout * sizeof(double), hipMemcpyDeviceToHost, s));
  hipError_t se = hipStreamSynchronize(s);
  double ms = 0;
  for (size_t i = 0; i + 1 < ev.size(); i += 2) {
    float x = 0.f;
    if (se == hipSuccess && hipEventElapsedTime(&x, ev[i], ev[i + 1]) == hipSuccess) ms += x;
  }
  for (auto e : ev) (void)hipEventDestroy(e);
  HIP_CHECK(se);
  c->last_items = dp->items;
  c->last_macs = dp->macs;
  c->last_ms = ms;
  c->last_variant = dp->variant;
}

}  // namespace xca

// ================================================================================== C ABI

extern "C" int64_t xca_num_correlator_backbones(xca_ctx *c, int m) {
  if (!c || m < 1) return -1;
  return ipow(2, m - 1) * ipow(c->M.P.n, m - 1) * ipow(c->M.P.p, m - 1);
}

extern "C" int xca_spgf(xca_ctx *c, const double *G, int m, const int *topo, const int32_t *fix, int64_t nfix, double *out) {
  try {
    if (!c) throw std::invalid_argument("xca: null context");
    HIP_CHECK(hipSetDevice(c->device));
    corr_eval(c, G, m, topo, fix, nfix, spgf_ops(c->M), true, out);
    return 0;
  } catch (std::exception &e) {
    xca::set_error(e.what());
    return -1;
  }
}

extern "C" int xca_correlator(xca_ctx *c, const double *G, int m, const int *topo, const int32_t *fix, int64_t nfix, int nmu,
                              const int *mu_conn, const double *mu_mats, int nkap, const int *kap_conn, const double *kap_mats, double *out) {
  try {
    if (!c) throw std::invalid_argument("xca: null context");
    HIP_CHECK(hipSetDevice(c->device));
    CorrOps co;
    co.nmu = nmu;
    co.nkap = nkap;
    add_ops(c->M.P, nmu, mu_conn, mu_mats, co.mu_conn, co.mu_off, co.mats, "O_tau");
    add_ops(c->M.P, nkap, kap_conn, kap_mats, co.kap_conn, co.kap_off, co.mats, "O_0");
    corr_eval(c, G, m, topo, fix, nfix, co, false, out);
    return 0;
  } catch (std::exception &e) {
    xca::set_error(e.what());
    return -1;
  }
}

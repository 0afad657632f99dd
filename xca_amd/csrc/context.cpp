// Host orchestration of the device path and the C ABI of the evaluator context.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/xca_amd.h"
#include "ctx.hpp"
#include "engine.hpp"
#include "host/errors.hpp"
#include "host/linalg.hpp"
#include "host/model.hpp"
#include "plan.hpp"
#include "ppsc.hpp"

namespace xca {

// tiles per chunk: enough chunks to cover the GPU several times, at most 64 tiles per chunk
// (XCA_CHUNK_TILES overrides: longer chunks reuse a shared prefix across more tiles)
// Queue order of a plan's chunks: the costliest first (columns in use summed over the chunk's
// tiles), so the cheap ones, such as programs' last partly filled tiles, fill the last round of the
// persistent launch instead of extending it.
void costliest_first(std::vector<DChunk> &chunks, const std::vector<DProg> &progs) {
  auto cost = [&](const DChunk &ch) {
    const DProg &d = progs[ch.prog];
    const int64_t items = std::min<int64_t>((int64_t)ch.t1 * d.items_per_tile, d.ntup) - (int64_t)ch.t0 * d.items_per_tile;
    return items * d.max_w;
  };
  std::stable_sort(chunks.begin(), chunks.end(), [&](const DChunk &a, const DChunk &b) { return cost(a) > cost(b); });
}

int chunk_tiles(int64_t total_tiles) {
  if (const char *e = std::getenv("XCA_CHUNK_TILES")) {
    const int v = std::atoi(e);
    if (v > 0) return v;
  }
  return (int)std::max<int64_t>(1, std::min<int64_t>(64, total_tiles / 2048));
}

}  // namespace xca

using namespace xca;


namespace {

Problem problem_from(const xca_problem *d) {
  if (!d) throw std::invalid_argument("xca_ctx_create: null problem");
  if (d->r <= 0 || d->p <= 0 || d->n <= 0 || d->nblocks <= 0) throw std::invalid_argument("xca_ctx_create: empty problem");
  if (d->coeffs_shape) {  // the extent check of the Fq-only constructor (block_sparse_backbone.cpp:28-40), before coeffs is read
    const int64_t *s = d->coeffs_shape;
    if (s[0] != d->p || s[1] != d->n || s[2] != d->n)
      throw std::invalid_argument("DiagramEvaluator: hyb_coeffs must have shape (p, n, n) with p = hyb_poles.size() = " + std::to_string(d->p) +
                                  " and n = sum(Fq.sym_set_sizes) = " + std::to_string(d->n) + ", got (" + std::to_string(s[0]) + ", " +
                                  std::to_string(s[1]) + ", " + std::to_string(s[2]) + ")");
  }
  Problem P;
  P.beta = d->beta; P.r = d->r; P.p = d->p; P.n = d->n; P.nb = d->nblocks;
  P.it.assign(d->it, d->it + P.r);
  P.rf.assign(d->rf, d->rf + P.r);
  P.poles.assign(d->poles, d->poles + P.p);
  P.coeffs.assign(d->coeffs, d->coeffs + (size_t)P.p * P.n * P.n);
  P.dims.assign(d->dims, d->dims + P.nb);
  P.c_conn.assign(d->c_conn, d->c_conn + (size_t)P.n * P.nb);
  P.cdag_conn.assign(d->cdag_conn, d->cdag_conn + (size_t)P.n * P.nb);
  auto unpack = [&](const std::vector<int> &conn, const double *s, std::vector<std::vector<double>> &dst) {
    dst.assign((size_t)P.n * P.nb, {});
    for (int o = 0; o < P.n; ++o)
      for (int j = 0; j < P.nb; ++j) {
        int t = conn[(size_t)o * P.nb + j];
        if (t < 0) continue;
        if (t >= P.nb) throw std::invalid_argument("xca_ctx_create: connection index out of range");
        size_t sz = (size_t)P.dims[t] * P.dims[j];
        dst[(size_t)o * P.nb + j].assign(s, s + sz);
        s += sz;
      }
  };
  unpack(P.c_conn, d->c_mats, P.c_mats);
  unpack(P.cdag_conn, d->cdag_mats, P.cdag_mats);
  return P;
}

// DLR operators as MFMA A-operand fragments: frag[(rb * KS + ks) * 64 + lane] = A[16 rb + (lane & 15)][4 ks + (lane >> 4)]
void pack_frag(const std::vector<double> &A, int rows, int cols, int RB, int KS, std::vector<double> &out) {
  for (int rb = 0; rb < RB; ++rb)
    for (int ks = 0; ks < KS; ++ks)
      for (int lane = 0; lane < 64; ++lane) {
        int i = 16 * rb + (lane & 15), k = 4 * ks + (lane >> 4);
        out.push_back((i < rows && k < cols) ? A[(size_t)i * cols + k] : 0.0);
      }
}

void build_frags(xca_ctx *c) {
  const auto &o = c->M.ops;
  const int r = o.r, RB = c->RB, RP = c->RP, KS = RP / 4;
  const double beta = c->M.P.beta;
  // L^{-1} P (explicit, unit lower factor only)
  std::vector<double> LiP((size_t)RP * RP, 0.0);
  {
    std::vector<double> e((size_t)r * r, 0.0);
    for (int i = 0; i < r; ++i) e[(size_t)i * r + i] = 1.0;
    for (int k = 0; k < r; ++k)
      if (o.piv[k] != k)
        for (int j = 0; j < r; ++j) std::swap(e[(size_t)k * r + j], e[(size_t)o.piv[k] * r + j]);
    for (int i = 1; i < r; ++i)
      for (int k = 0; k < i; ++k) {
        double l = o.lu[(size_t)i * r + k];
        if (l != 0.0)
          for (int j = 0; j < r; ++j) e[(size_t)i * r + j] -= l * e[(size_t)k * r + j];
      }
    for (int i = 0; i < r; ++i)
      for (int j = 0; j < r; ++j) LiP[(size_t)i * RP + j] = e[(size_t)i * r + j];
  }
  std::vector<double> Up((size_t)RP * RP, 0.0);
  for (int i = 0; i < RP; ++i)
    for (int k = i; k < RP; ++k) Up[(size_t)i * RP + k] = (i < r && k < r) ? o.lu[(size_t)i * r + k] : (i == k ? 1.0 : 0.0);
  std::vector<double> Wp((size_t)RP * RP, 0.0), M2((size_t)RP * 2 * RP, 0.0);
  for (int i = 0; i < r; ++i)
    for (int k = 0; k < r; ++k) {
      Wp[(size_t)i * RP + k] = o.W[(size_t)i * r + k];
      M2[(size_t)i * 2 * RP + k] = beta * o.cf2it[(size_t)i * r + k];
      M2[(size_t)i * 2 * RP + RP + k] = beta * o.tcf2it[(size_t)i * r + k];
    }
  std::vector<double> Ep((size_t)RP * RP, 0.0);
  for (int i = 0; i < r; ++i)
    for (int k = 0; k < r; ++k) Ep[(size_t)i * RP + k] = dlr::kernel(-o.it[i], o.rf[k]);  // coefficients -> values at -t_i
  std::vector<double> fLiP, fW, fM2, fAU, fD, fE;
  pack_frag(Ep, RP, RP, RB, KS, fE);
  pack_frag(LiP, RP, RP, RB, KS, fLiP);
  pack_frag(Wp, RP, RP, RB, KS, fW);
  pack_frag(M2, RP, 2 * RP, RB, 2 * KS, fM2);
  // 4x4 diagonal-block inverses and trailing-update fragments
  const int S = RP / 4;
  fD.assign((size_t)S * 16, 0.0);
  for (int s = 0; s < S; ++s) {
    for (int cidx = 0; cidx < 4; ++cidx) {  // inverse column by back substitution
      double x[4] = {0, 0, 0, 0};
      x[cidx] = 1.0;
      for (int i = 3; i >= 0; --i) {
        for (int k = i + 1; k < 4; ++k) x[i] -= Up[(size_t)(4 * s + i) * RP + 4 * s + k] * x[k];
        x[i] /= Up[(size_t)(4 * s + i) * RP + 4 * s + i];
      }
      for (int i = 0; i < 4; ++i) fD[(size_t)s * 16 + i * 4 + cidx] = x[i];
    }
  }
  fAU.assign((size_t)S * RB * 64, 0.0);
  for (int s = 0; s < S; ++s)
    for (int rb = 0; rb <= (s >> 2); ++rb)
      for (int lane = 0; lane < 64; ++lane) {
        int i = 16 * rb + (lane & 15), k = 4 * s + (lane >> 4);
        fAU[((size_t)s * RB + rb) * 64 + lane] = (i < 4 * s) ? -Up[(size_t)i * RP + k] : 0.0;
      }
  std::vector<double> all;
  size_t oLiP = 0;
  all.insert(all.end(), fLiP.begin(), fLiP.end());
  size_t oW = all.size();
  all.insert(all.end(), fW.begin(), fW.end());
  size_t oM2 = all.size();
  all.insert(all.end(), fM2.begin(), fM2.end());
  size_t oAU = all.size();
  all.insert(all.end(), fAU.begin(), fAU.end());
  size_t oD = all.size();
  all.insert(all.end(), fD.begin(), fD.end());
  // -U strictly above the diagonal, packed by rows
  size_t oU = all.size();
  for (int i = 0; i < RP; ++i)
    for (int k = i + 1; k < RP; ++k) all.push_back(-Up[(size_t)i * RP + k]);
  size_t oE = all.size();
  all.insert(all.end(), fE.begin(), fE.end());
  c->frag.upload(all, c->stream);
  c->D.E = c->frag.ptr + oE;
  c->D.LiP = c->frag.ptr + oLiP;
  c->D.W = c->frag.ptr + oW;
  c->D.M2 = c->frag.ptr + oM2;
  c->D.AU = c->frag.ptr + oAU;
  c->D.Dinv = c->frag.ptr + oD;
  c->D.Upk = c->frag.ptr + oU;
  c->D.r = r;
  c->D.RP = RP;
}

int64_t ipow(int64_t a, int e) {
  int64_t r = 1;
  for (int i = 0; i < e; ++i) r *= a;
  return r;
}

// Program template of one block path: ops resolved to device form, its pole-tuple count and (for an
// explicit flat-index subset) the requested tuples.
struct ProgSpec {
  DProg d;
  std::vector<int> tuples;  // explicit subset: tuple indices in the class enumeration
  bool explicit_subset = false;
  int tup_off = -1;         // its offset in the plan's tuple list
  int nl = 0;
  int fb[kMaxLines + 1];
  int cls = 0;
};

}  // namespace

namespace xca {

// canonical multiply-add count of one (path, pole tuple) evaluation; a reflection is priced as
// vals2coefs plus the evaluation at the reflected nodes (2 r^2 per column)
double canonical_macs(const Program &pg, int r) {
  double s = pg.refl ? 2.0 * r * r * pg.out_rows * pg.out_cols : 0.0;
  for (const auto &op : pg.ops) {
    if (op.type == OP_V) s += (double)r * op.bout * op.bin * op.cols;
    else if (op.type == OP_E) s += 4.0 * r * r * op.bin * op.cols + 3.0 * r * op.bin * op.bin * op.cols;
    else s += (double)op.qs * r * op.bin * op.b1 * op.cols + (double)op.qs * op.qs * r * op.bin * op.cols + (double)op.qs * r * op.bout * op.bin * op.cols;
  }
  return s;
}

// device form of one op
DOp device_op(const Op &op, const std::vector<EdgeTable> &tables, const xca_ctx *c) {
  const Problem &P = c->M.P;
  DOp o{};
  o.type = op.type; o.bin = op.bin; o.bout = op.bout; o.cols = op.cols;
  o.fbase = op.fbase; o.fstride_l = op.fstride_l;
  o.bar_line = op.bar_line; o.kx = op.kx; o.ks = op.ks; o.nlines = op.nlines;
  for (int q = 0; q < op.nlines; ++q) o.lines[q] = op.lines[q];
  if (op.type == OP_E) {
    const auto &tb = tables[op.etab];
    int64_t esz = 2 * (int64_t)P.r * P.dims[tb.block] * P.dims[tb.block];
    o.ebase = tb.offset;
    for (int q = 0; q < op.nlines; ++q) o.estride[q] = ipow(P.p, q) * esz;
  }
  o.mubase = op.mubase; o.qs = op.qs; o.forward = op.forward; o.b1 = op.b1; o.set = op.set;
  if (op.type == OP_Z) {
    // the device reads the tabulated kappa sums Fh (r, qs, b1, b0) in place of the kappa operators
    if (op.fhbase < 0) throw std::logic_error("xca: zero-vertex kappa sums are not tabulated for this problem");
    o.fbase = op.fhbase;
    for (int k = 0; k < op.qs; ++k) o.orb[k] = c->M.S.to_orb[op.set][k];
  }
  return o;
}

// edge-table jobs: entries of each table in runs of at most 32 columns (64 for blocks above 4
// states, one entry for blocks above 8), so that a table launch spreads over many workgroups
std::vector<DTabJob> tab_jobs(const std::vector<EdgeTable> &tables, const xca_ctx *c) {
  const Problem &P = c->M.P;
  std::vector<DTabJob> out;
  for (const auto &tb : tables) {
    int b = P.dims[tb.block], bb = b * b;
    if (bb > tile_columns(0)) throw std::invalid_argument("xca: block too large for the edge-table kernel");
    int per = std::max(1, (bb <= 16 ? 16 : bb <= 32 ? 32 : 64) / bb);
    for (int64_t e0 = 0; e0 < tb.nentries; e0 += per) {
      DTabJob j{};
      j.block = tb.block; j.b = b; j.nlines = tb.nlines; j.e0 = (int)e0;
      j.ne = (int)std::min<int64_t>(per, tb.nentries - e0);
      j.g_off = c->M.G_off[tb.block];
      j.tr = tb.tr;
      for (int q = 0; q < tb.nlines; ++q) j.lsign[q] = tb.lsign[q];
      j.out_off = tb.offset;
      out.push_back(j);
    }
  }
  return out;
}

static int64_t variant_bytes(const DevVariant &V) {
  return (int64_t)(V.progs.n * sizeof(DProg) + V.chunks.n * sizeof(DChunk) + V.slot.n * sizeof(int64_t) + V.red_chunks.n * sizeof(int) +
                   V.red_range.n * sizeof(I2) + V.partial.n * sizeof(double));
}

// Mark the context's current zero-block pattern as most recently used. When more than
// kMaxZeroPatterns patterns have plans, the plans of the least recently used one are released: the
// combined-launch descriptors and every device launch that may use them are finished first.
static void touch_zero_pattern(xca_ctx *c) {
  auto &L = c->zero_lru;
  auto f = std::find(L.begin(), L.end(), c->G_zero);
  if (f != L.end()) L.erase(f);
  L.push_back(c->G_zero);
  while ((int)L.size() > xca_ctx::kMaxZeroPatterns) {
    const std::vector<int> old = L.front();
    L.erase(L.begin());
    HIP_CHECK(hipDeviceSynchronize());
    c->sigall.reset();
    const size_t nb = old.size();
    for (auto e = c->plans.begin(); e != c->plans.end();) {
      const auto &k = e->first;
      const size_t z0 = 1 + 2 * (size_t)k[0];
      if (k.size() >= z0 + nb && std::equal(old.begin(), old.end(), k.begin() + z0)) e = c->plans.erase(e);
      else ++e;
    }
  }
}

DevPlan *get_plan(xca_ctx *c, int m, const int *topo, const int32_t *fix, int64_t nfix) {
  std::vector<int> key;
  key.push_back(m);
  key.insert(key.end(), topo, topo + 2 * m);
  key.insert(key.end(), c->G_zero.begin(), c->G_zero.end());
  if (fix) {
    key.push_back(-1);
    key.insert(key.end(), fix, fix + nfix);
  }
  touch_zero_pattern(c);
  auto it = c->plans.find(key);
  if (it != c->plans.end()) return it->second.get();
  if (fix)  // one explicit-subset plan at a time: a caller sweeping flat indices does not accumulate plans
    for (auto e = c->plans.begin(); e != c->plans.end();) {
      if (std::find(e->first.begin(), e->first.end(), -1) == e->first.end()) {
        ++e;
        continue;
      }
      HIP_CHECK(hipStreamSynchronize(c->stream));  // its launches may still be in flight
      e = c->plans.erase(e);
    }

  const Problem &P = c->M.P;
  const int r = P.r, p = P.p, n = P.n;
  if (m < 1 || m > kMaxLines + 1) throw std::invalid_argument("xca: diagram order must be between 1 and 4");
  Plan plan;
  build_programs(c->M, c->ps, topo, m, 0, c->G_zero, plan);
  auto dp = std::make_unique<DevPlan>();
  const int nl = m - 1;
  // explicit tuple lists when a subset of flat indices is requested
  std::map<std::pair<int, int>, std::vector<int>> wanted;  // (fb_ix, o_ix) -> p_ix list
  const int64_t o_max = ipow(n, m - 1), p_max = ipow(p, m - 1), f_max = ipow(2, m) * o_max * p_max;
  if (fix) {
    for (int64_t i = 0; i < nfix; ++i) {
      int64_t f = fix[i];
      if (f < 0 || f >= f_max) throw std::runtime_error("DiagramEvaluator::eval_self_energy: f_ix out of range");
      int o_ix = (int)(f % o_max);
      int64_t rest = f / o_max;
      wanted[{(int)(rest / p_max), o_ix}].push_back((int)(rest % p_max));
    }
  }
  // program specs (tile-width independent)
  std::vector<ProgSpec> specs;
  for (auto &pg : plan.progs) {
    for (auto &op : pg.ops) {
      if (op.bin > 12 || op.bout > 12 || (op.type == OP_Z && op.qs > kMaxSet))
        throw std::invalid_argument("xca: device path supports blocks of dimension <= 12 and symmetry sets of size <= 8");
    }
    if (pg.max_w > tile_columns(0)) throw std::invalid_argument("xca: block path too wide for one tile");
    ProgSpec sp;
    DProg &d = sp.d;
    d = DProg{};
    d.op0 = (int)dp->h_ops.size();
    d.nops = (int)pg.ops.size();
    d.sigma_block = pg.sigma_block;
    d.b0 = pg.b0;
    d.b1 = pg.b1;
    d.g_init = c->M.G_off[pg.init_block];
    d.sign = pg.sign;
    d.nlines = nl;
    d.max_w = pg.max_w;
    d.gtr = 0;
    d.out_off = -1;
    sp.nl = nl;
    sp.cls = pg.cls;
    for (int i = 0, f = pg.fb_ix; i < m; ++i, f /= 2) sp.fb[i] = f % 2;
    int64_t ntup = 1;
    for (int x = 0; x < nl; ++x) {
      d.pk_sign[x] = pg.pk_sign[x];
      d.pk_exp[x] = pg.pk_exp[x];
      int fwd = sp.fb[x + 1], rb = (pg.cls >> x) & 1;
      d.set_off[x] = c->plist_off[(fwd ? 0 : 2) + rb];
      d.set_len[x] = (int)c->ps.get(fwd, rb).size();
      ntup *= d.set_len[x];
    }
    if (fix) {
      sp.explicit_subset = true;
      auto w = wanted.find({pg.fb_ix, pg.o_ix});
      if (w == wanted.end()) continue;
      for (int pix : w->second) {
        // the pole index runs with line 0 fastest (backbone.cpp:141-150); the device's tuple index
        // with line 0 slowest, in the order the chain kernel decodes it
        int u = 0, cls = 0, pp = pix;
        for (int x = 0; x < nl; ++x) {
          int l = pp % p;
          pp /= p;
          int fwd = sp.fb[x + 1];
          double wl = P.poles[l] * P.beta;
          int rb = fwd ? (wl <= 0 ? 0 : 1) : (wl >= 0 ? 0 : 1);
          cls |= rb << x;
          const auto &lst = c->ps.get(fwd, rb);
          int pos = (int)(std::find(lst.begin(), lst.end(), l) - lst.begin());
          u = u * (int)lst.size() + pos;
        }
        if (cls == pg.cls) sp.tuples.push_back(u);
      }
      if (sp.tuples.empty()) continue;
      ntup = (int64_t)sp.tuples.size();
      sp.tup_off = (int)dp->h_tuples.size();
      dp->h_tuples.insert(dp->h_tuples.end(), sp.tuples.begin(), sp.tuples.end());
    }
    d.ntup = (int)ntup;
    // ops before the first one that involves a line other than the first depend on the first
    // line's pole only; items sharing that pole share them (not for explicit subsets)
    d.split = 0;
    d.share_len = 1;
    if (nl >= 2 && !fix) {
      int split = (int)pg.ops.size();
      for (size_t i = 0; i < pg.ops.size() && split == (int)pg.ops.size(); ++i) {
        const auto &op = pg.ops[i];
        bool other = false;
        if (op.type == OP_V) other = op.bar_line >= 1 || (op.ks != 0 && op.kx >= 1);
        else if (op.type == OP_E)
          for (int q = 0; q < op.nlines; ++q) other = other || op.lines[q] >= 1;
        if (other) split = (int)i;
      }
      if (split > 0 && split < (int)pg.ops.size()) {
        d.split = split;
        d.share_len = (int)(ntup / d.set_len[0]);
      }
    }
    for (auto &op : pg.ops) dp->h_ops.push_back(device_op(op, plan.tables, c));
    dp->items += ntup;
    dp->macs += canonical_macs(pg, r) * (double)ntup;
    specs.push_back(std::move(sp));
  }
  hipStream_t s = c->stream;
  // one variant per tile width
  // candidates: every tile width, or the forced one with the widest as fallback
  std::vector<int> cand;
  if (c->tile_variant >= 0) {
    cand.push_back(c->tile_variant);
    if (c->tile_variant != 0) cand.push_back(0);
  } else {
    for (int v = 0; v < num_tile_variants(); ++v) cand.push_back(v);
  }
  dp->var.resize(num_tile_variants());
  for (int v : cand) {
    const int NT = tile_columns(v);
    bool ok = true;
    int64_t est_tiles = 0;
    for (auto &sp : specs) {
      if (sp.d.max_w > NT) { ok = false; break; }
      int per = std::max(1, NT / sp.d.max_w);
      est_tiles += (sp.d.ntup + per - 1) / per;
    }
    if (!ok) continue;
    auto V = std::make_unique<DevVariant>();
    V->NT = NT;
    const int ct = chunk_tiles(est_tiles);
    for (auto &sp : specs) {
      DProg d = sp.d;
      d.items_per_tile = std::max(1, NT / d.max_w);
      if (!sp.explicit_subset || c->launch_mode == 0) {
        // an explicit subset keeps its requested tuples in a list that the items index
        if (sp.explicit_subset) d.tup_off = sp.tup_off;
        int pidx = (int)V->h_progs.size();
        V->h_progs.push_back(d);
        int ntiles = (d.ntup + d.items_per_tile - 1) / d.items_per_tile;
        for (int t0 = 0; t0 < ntiles; t0 += ct) V->h_chunks.push_back(DChunk{pidx, t0, std::min(ntiles, t0 + ct), 0, 0});
      } else {
        // launch mode 1: each requested tuple becomes a one-tuple program whose pole lists have
        // length 1, so that one chunk is one (path, pole tuple) item
        for (int u : sp.tuples) {
          DProg d1 = d;
          int decode = u;
          for (int x = sp.nl - 1; x >= 0; --x) {  // line nl - 1 is the fastest digit
            int len = d.set_len[x];
            int pos = decode % len;
            decode /= len;
            d1.set_off[x] = c->plist_off[(sp.fb[x + 1] ? 0 : 2) + ((sp.cls >> x) & 1)] + pos;
            d1.set_len[x] = 1;
          }
          d1.ntup = 1;
          int pidx = (int)V->h_progs.size();
          V->h_progs.push_back(d1);
          V->h_chunks.push_back(DChunk{pidx, 0, 1, 0, 0});
        }
      }
    }
    // the per-block reduction lists, and the partial slots laid out block by block in reduction
    // order: the slots of one Sigma block are contiguous (slot k of the block's list at its first
    // slot + k r b0^2), which is how the reduction kernels address them
    int64_t off = 0;
    std::vector<std::vector<int>> per_block(P.nb);
    V->h_slot.resize(V->h_chunks.size());
    for (size_t ci = 0; ci < V->h_chunks.size(); ++ci) per_block[V->h_progs[V->h_chunks[ci].prog].sigma_block].push_back((int)ci);
    for (int b = 0; b < P.nb; ++b) {
      I2 rg;
      rg.x = (int)V->h_red.size();
      for (int ci : per_block[b]) {
        const DProg &d = V->h_progs[V->h_chunks[ci].prog];
        if (d.b0 != P.dims[b]) throw std::logic_error("xca: program block dimension differs from its Sigma block");
        V->h_chunks[ci].slot = off;
        V->h_slot[ci] = off;
        off += (int64_t)r * d.b0 * d.b0;
        V->h_red.push_back(ci);
      }
      rg.y = (int)V->h_red.size();
      V->h_red_range.push_back(rg);
    }
    // each chunk keeps its slot and the reduction its order, so the queue order does not change
    // the result
    costliest_first(V->h_chunks, V->h_progs);
    V->nchunks = (int)V->h_chunks.size();
    V->progs.upload(V->h_progs, s);
    V->chunks.upload(V->h_chunks, s);
    V->slot.upload(V->h_slot, s);
    V->red_chunks.upload(V->h_red, s);
    V->red_range.upload(V->h_red_range, s);
    V->partial.alloc(std::max<int64_t>(off, 1));
    dp->var[v] = std::move(V);
    if (c->tile_variant >= 0 && dp->chosen < 0) dp->chosen = v;
  }
  dp->h_tab = tab_jobs(plan.tables, c);
  dp->ops.upload(dp->h_ops, s);
  dp->tuples.upload(dp->h_tuples.empty() ? std::vector<int>{0} : dp->h_tuples, s);
  dp->tabjobs.upload(dp->h_tab, s);
  dp->ftab.alloc(std::max<int64_t>(plan.ftab_size, 1));
  dp->queue.alloc(2);
  HIP_CHECK(init_queue(dp->queue.ptr, s));
  dp->ntabjobs = (int)dp->h_tab.size();
  dp->device_bytes = (int64_t)(dp->ops.n * sizeof(DOp) + dp->tabjobs.n * sizeof(DTabJob) + dp->ftab.n * sizeof(double) + dp->tuples.n * sizeof(int));
  for (auto &V : dp->var)
    if (V) dp->device_bytes += variant_bytes(*V);
  HIP_CHECK(hipStreamSynchronize(s));
  DevPlan *raw = dp.get();
  c->plans[key] = std::move(dp);
  return raw;
}

ChainArgs chain_args(xca_ctx *c, DevPlan *dp, DevVariant *V) {
  ChainArgs A{};
  A.D = c->D;
  A.ops = dp->ops.ptr; A.progs = V->progs.ptr; A.chunks = V->chunks.ptr;
  A.plist = c->plist.ptr; A.G = c->G.ptr; A.opmats = c->opmats.ptr; A.ftab = dp->ftab.ptr; A.kp = c->kp.ptr; A.km = c->km.ptr;
  A.kpow = c->kpow.ptr; A.partial = V->partial.ptr; A.tuples = dp->tuples.ptr;
  A.nchunks = V->nchunks;
  A.r = c->M.P.r; A.p = c->M.P.p; A.maxexp = c->maxexp;
  A.maxb = ctx_max_block(c);
  return A;
}

int ctx_max_block(const xca_ctx *c) {
  int m = 1;
  for (int b = 0; b < c->M.P.nb; ++b) m = std::max(m, c->M.P.dims[b]);
  return m;
}

// Tile-width autotuner: time every applicable variant on this plan (all chunks when the plan is
// small, else an evenly strided sample of about 2048 chunks, scaled by the items it covers) and
// keep the fastest. Runs once per plan; the partial slots written here are rewritten by the real
// launch (the first tile of a chunk overwrites its slot).
void autotune(xca_ctx *c, DevPlan *dp) {
  hipStream_t s = c->stream;
  int best = -1;
  for (int v = 0; v < (int)dp->var.size(); ++v) {
    DevVariant *V = dp->var[v].get();
    if (!V || V->nchunks == 0) continue;
    const int stride = std::max(1, V->nchunks / 2048);
    DevBuf<DChunk> sample;
    int64_t items_all = 0, items_sample = 0;
    std::vector<DChunk> hs;
    for (int ci = 0; ci < V->nchunks; ++ci) {
      const DChunk &ch = V->h_chunks[ci];
      const DProg &d = V->h_progs[ch.prog];
      int64_t it = std::min<int64_t>((int64_t)ch.t1 * d.items_per_tile, d.ntup) - (int64_t)ch.t0 * d.items_per_tile;
      items_all += it;
      if (ci % stride == 0) {
        hs.push_back(ch);
        items_sample += it;
      }
    }
    std::vector<ChainArgs> A(1, chain_args(c, dp, V));
    if (stride > 1) {
      sample.upload(hs, s);
      A[0].chunks = sample.ptr;
      A[0].nchunks = (int)hs.size();
    }
    std::vector<int> pre{0, A[0].nchunks};
    DevBuf<ChainArgs> dA;
    DevBuf<int> dpre;
    dA.upload(A, s);
    dpre.upload(pre, s);
    DevBuf<double> scr;
    const int64_t sstride = (int64_t)c->RP * tile_columns(v);
    scr.alloc((size_t)chain_resident(c->RB, v) * sstride);
    ChainBatch B{dA.ptr, dpre.ptr, 1, A[0].nchunks, dp->queue.ptr, scr.ptr, sstride};
    B.maxb = A[0].maxb;
    float best_ms = 1e30f;
    for (int rep = 0; rep < 2; ++rep) {
      HIP_CHECK(hipEventRecord(c->ev0, s));
      HIP_CHECK(launch_chain(c->RB, v, B, s));
      HIP_CHECK(hipEventRecord(c->ev1, s));
      HIP_CHECK(hipEventSynchronize(c->ev1));
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      best_ms = std::min(best_ms, ms);
    }
    V->ms = best_ms * (double)items_all / (double)std::max<int64_t>(items_sample, 1);
    if (best < 0 || V->ms < dp->var[best]->ms) best = v;
  }
  if (best < 0) throw std::runtime_error("xca: no tile variant applies to this plan");
  HIP_CHECK(hipStreamSynchronize(s));
  dp->chosen = best;
  // release the other variants' buffers, except small ones (a combined launch of several plans
  // may run a small plan at the tile width of a larger one)
  for (int v = 0; v < (int)dp->var.size(); ++v)
    if (v != best && dp->var[v] && variant_bytes(*dp->var[v]) > (16 << 20)) {
      dp->device_bytes -= variant_bytes(*dp->var[v]);
      dp->var[v].reset();
    }
}

TabSeg tab_seg(xca_ctx *c, DevPlan *dp) {
  TabSeg T{};
  T.D = c->D;
  T.jobs = dp->tabjobs.ptr;
  T.G = c->G.ptr;
  T.kp = c->kp.ptr;
  T.km = c->km.ptr;
  T.ftab = dp->ftab.ptr;
  T.p = c->M.P.p;
  return T;
}

void accumulate(xca_ctx *c, int m, const int *topo, const int32_t *fix, int64_t nfix, double scale) {
  if (!c->have_G) throw std::runtime_error("xca: G has not been set");
  if (c->wide) {
    wide_accumulate(c, m, topo, fix, nfix, scale);
    return;
  }
  DevPlan *dp = get_plan(c, m, topo, fix, nfix);
  hipStream_t s = c->stream;
  c->last_items = dp->items;
  c->last_macs = dp->macs;
  c->last_ms = 0;
  if (dp->items == 0) return;
  if (dp->ntabjobs > 0) {
    if (dp->h_tseg.empty()) {
      dp->h_tseg.assign(1, tab_seg(c, dp));
      dp->h_tseg_prefix = {0, dp->ntabjobs};
      dp->tseg.upload(dp->h_tseg, s);
      dp->tseg_prefix.upload(dp->h_tseg_prefix, s);
    }
    HIP_CHECK(launch_tab(c->RB, TabBatch{dp->tseg.ptr, dp->tseg_prefix.ptr, 1, dp->ntabjobs}, s));
  }
  if (dp->chosen < 0) autotune(c, dp);
  DevVariant *V = dp->var[dp->chosen].get();
  c->last_variant = dp->chosen;
  if (V->h_seg.empty()) {
    V->h_seg.assign(1, chain_args(c, dp, V));
    V->h_seg_prefix = {0, V->nchunks};
    V->seg.upload(V->h_seg, s);
    V->seg_prefix.upload(V->h_seg_prefix, s);
  }
  const int64_t stride = (int64_t)c->RP * V->NT;
  if (!V->scratch.ptr) V->scratch.alloc((size_t)chain_resident(c->RB, dp->chosen) * stride);
  if (c->launch_mode == 1 && !V->item_seg.ptr) {
    std::vector<ChainArgs> a(V->nchunks, V->h_seg[0]);
    for (int k = 0; k < V->nchunks; ++k) {
      a[k].chunks = V->chunks.ptr + k;
      a[k].nchunks = 1;
    }
    V->item_seg.upload(a, s);
    V->item_prefix.upload(std::vector<int>{0, 1}, s);
  }
  HIP_CHECK(hipEventRecord(c->ev0, s));
  if (c->launch_mode == 1) {
    // one launch per chunk; an explicit flat-index subset has one chunk per (path, pole tuple) item
    for (int k = 0; k < V->nchunks; ++k) {
      ChainBatch B{V->item_seg.ptr + k, V->item_prefix.ptr, 1, 1, dp->queue.ptr, V->scratch.ptr, stride};
      B.maxb = V->h_seg[0].maxb;
      HIP_CHECK(launch_chain(c->RB, dp->chosen, B, s));
    }
  } else {
    ChainBatch B{V->seg.ptr, V->seg_prefix.ptr, 1, V->nchunks, dp->queue.ptr, V->scratch.ptr, stride};
    B.maxb = V->h_seg[0].maxb;
    HIP_CHECK(launch_chain(c->RB, dp->chosen, B, s));
  }
  HIP_CHECK(hipEventRecord(c->ev1, s));
  int maxsz = 0;
  for (int b = 0; b < c->M.P.nb; ++b) maxsz = std::max(maxsz, c->M.P.r * c->M.P.dims[b] * c->M.P.dims[b]);
  HIP_CHECK(launch_reduce(V->slot.ptr, V->red_chunks.ptr, V->red_range.ptr, c->sig_off.ptr, c->sig_size.ptr, c->M.P.nb, maxsz,
                          V->partial.ptr, c->sigma.ptr, scale, s));
}

// Connected pairings of 2*order vertices with their parity sign (python/triqs_xca/diag.py:48-127):
// vertex 0 is paired first with each later vertex, the sign flips for an even distance, and a
// pairing is kept when every line is reachable through crossings from the line attached to vertex 0.
static void pairings(int parity, std::vector<std::pair<int, int>> &cur, std::vector<int> rest,
                     std::vector<std::pair<int, std::vector<std::pair<int, int>>>> &out) {
  if (rest.empty()) {
    out.push_back({parity, cur});
    return;
  }
  for (size_t i = 1; i < rest.size(); ++i) {
    std::vector<int> r2;
    for (size_t k = 1; k < rest.size(); ++k)
      if (k != i) r2.push_back(rest[k]);
    cur.push_back({rest[0], rest[i]});
    pairings(i % 2 == 0 ? -parity : parity, cur, r2, out);
    cur.pop_back();
  }
}

static bool connected(const std::vector<std::pair<int, int>> &pr) {
  auto cross = [](std::pair<int, int> a, std::pair<int, int> b) {
    return (a.first < b.first && b.first < a.second && a.second < b.second) || (b.first < a.first && a.first < b.second && b.second < a.second);
  };
  std::vector<std::pair<int, int>> stack, rest;
  for (auto &p : pr) ((p.first <= 0) != (p.second <= 0) ? stack : rest).push_back(p);
  while (!stack.empty()) {
    auto pc = stack.front();
    stack.erase(stack.begin());
    for (int i = (int)rest.size() - 1; i >= 0; --i)
      if (cross(pc, rest[i])) {
        stack.push_back(rest[i]);
        rest.erase(rest.begin() + i);
      }
  }
  return rest.empty();
}

std::vector<std::pair<int, std::vector<int>>> connected_topologies(int order) {
  std::vector<std::pair<int, std::vector<std::pair<int, int>>>> all;
  std::vector<std::pair<int, int>> cur;
  std::vector<int> v(2 * order);
  for (int i = 0; i < 2 * order; ++i) v[i] = i;
  pairings(1, cur, v, all);
  std::vector<std::pair<int, std::vector<int>>> out;
  for (auto &pp : all)
    if (connected(pp.second)) {
      std::vector<int> t;
      for (auto &p : pp.second) { t.push_back(p.first); t.push_back(p.second); }
      out.push_back({pp.first, t});
    }
  return out;
}

// Sigma accumulator <- sum over orders 1..max_order and connected topologies with the solver's net
// factor (-1)^order * s * (-1)^(order+1) = -s (block_sparse_solver.py:609-638)
void sigma_full_order(xca_ctx *c, int max_order) {
  HIP_CHECK(launch_fill(c->M.G_size, 0.0, c->sigma.ptr, c->stream));
  for (int order = 1; order <= max_order; ++order)
    for (auto &st : connected_topologies(order)) accumulate(c, order, st.second.data(), nullptr, 0, -(double)st.first);
}

}  // namespace xca

// ================================================================================== C ABI

#define XCA_TRY(body)                  \
  try {                                \
    body                               \
  } catch (std::exception & e) {       \
    xca::set_error(e.what());          \
    return -1;                         \
  }

namespace xca {

// Host part of a context: the model, symmetry sets, operators and pole classes. No device memory.
xca_ctx *ctx_init_host(const xca_problem *desc) {
  auto c = std::make_unique<xca_ctx>();
  c->M.init(problem_from(desc));
  const Problem &P = c->M.P;
  std::vector<double> w(P.p);
  for (int l = 0; l < P.p; ++l) w[l] = P.poles[l] * P.beta;
  c->ps.init(w);
  // padded DLR rank: 32 or 64 rows (16-row MFMA blocks; the row-partitioned GEMMs need the row-block
  // count to divide the waves of a workgroup)
  if (P.r > 64) throw std::invalid_argument("xca: DLR rank above 64 is not supported by the device path");
  c->RB = (P.r <= 32) ? 2 : 4;
  c->RP = 16 * c->RB;
  if (const char *tv = std::getenv("XCA_TILE_VARIANT")) c->tile_variant = std::atoi(tv);
  // blocks above 12 rows go to the large-block engine (XCA_WIDE=1 forces it for any problem)
  c->wide = *std::max_element(P.dims.begin(), P.dims.end()) > 12;
  if (const char *w = std::getenv("XCA_WIDE")) c->wide = c->wide || std::atoi(w) != 0;
  if (c->tile_variant >= num_tile_variants()) c->tile_variant = -1;
  return c.release();
}

// Device part: stream, events, resident operator and kernel tables, DLR fragments.
void ctx_init_device(xca_ctx *c, int device) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw std::runtime_error("xca_ctx_create: no GPU device available");
  if (device < 0 || device >= ndev) throw std::runtime_error("xca_ctx_create: device index out of range");
  c->device = device;
  HIP_CHECK(hipSetDevice(device));
  const Problem &P = c->M.P;
  HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIP_CHECK(hipEventCreate(&c->ev0));
  HIP_CHECK(hipEventCreate(&c->ev1));
  hipStream_t s = c->stream;
  HIP_CHECK(prepare_kernels(c->RB, s));
  HIP_CHECK(launch_dyson(DysonBatch{}, s));  // the Dyson kernel's attribute, likewise
  // Record the timing events once before any graph capture on this stream: an event whose first
  // record comes after a capture on its stream is rejected as an invalid handle by the runtime
  // (a forced tile width skips the autotuner, which otherwise records them first).
  HIP_CHECK(hipEventRecord(c->ev0, s));
  HIP_CHECK(hipEventRecord(c->ev1, s));
  c->opmats.upload(c->M.opmats, s);
  c->kp.upload(c->M.kp, s);
  c->km.upload(c->M.km, s);
  c->delta.upload(c->M.delta, s);
  c->delta_refl.upload(c->M.delta_refl, s);
  std::vector<double> kpow((size_t)2 * c->maxexp * P.p);
  for (int sg = 0; sg < 2; ++sg)
    for (int e = 1; e <= c->maxexp; ++e)
      for (int l = 0; l < P.p; ++l) kpow[((size_t)sg * c->maxexp + (e - 1)) * P.p + l] = std::pow(sg == 0 ? c->M.k0p[l] : c->M.k0m[l], -e);
  c->kpow.upload(kpow, s);
  // pole lists
  const std::vector<int> *lists[4] = {&c->ps.fwd_a, &c->ps.fwd_b, &c->ps.bwd_a, &c->ps.bwd_b};
  c->plist_h.clear();
  for (int i = 0; i < 4; ++i) {
    c->plist_off[i] = (int)c->plist_h.size();
    c->plist_h.insert(c->plist_h.end(), lists[i]->begin(), lists[i]->end());
  }
  c->plist.upload(c->plist_h, s);
  std::vector<int64_t> so(P.nb);
  std::vector<int> ss(P.nb);
  for (int b = 0; b < P.nb; ++b) {
    so[b] = c->M.G_off[b];
    ss[b] = P.r * P.dims[b] * P.dims[b];
  }
  c->sig_off.upload(so, s);
  c->sig_size.upload(ss, s);
  c->G.alloc(c->M.G_size);
  c->sigma.alloc(c->M.G_size);
  HIP_CHECK(launch_fill(c->M.G_size, 0.0, c->sigma.ptr, s));
  build_frags(c);
  HIP_CHECK(hipStreamSynchronize(s));
}

// Device bytes of the resident context tables (without plans and Dyson state).
int64_t ctx_bytes(const xca_ctx *c) {
  const Problem &P = c->M.P;
  const int64_t RP = 16 * c->RB;
  return 8 * ((int64_t)c->M.opmats.size() + c->M.kp.size() + c->M.km.size() + c->M.delta.size() + c->M.delta_refl.size() +
              2 * c->maxexp * P.p + 2 * c->M.G_size + 13 * RP * RP) +
         4 * (4 * P.p) + 12 * P.nb;  // pole lists; Sigma block offsets (int64) and sizes (int)
}

// Upper bound of the device bytes a plan of this topology takes at the context's current zero-block
// pattern: edge tables, op list and table jobs once, plus, for every tile variant get_plan allocates
// (the forced one and the widest as its fallback, or all of them when the autotuner chooses), its
// programs, chunk descriptors, partial slots and reduction lists at that variant's tile count.
int64_t plan_bytes_estimate(xca_ctx *c, int m, const int *topo) {
  if (c->wide) return wide_plan_bytes_estimate(c, m, topo);  // Sigma on the large-block engine
  Plan plan;
  build_programs(c->M, c->ps, topo, m, 0, c->G_zero, plan);
  const Problem &P = c->M.P;
  std::vector<int> cand;
  if (c->tile_variant >= 0) {
    cand.push_back(c->tile_variant);
    if (c->tile_variant != 0) cand.push_back(0);
  } else {
    for (int v = 0; v < num_tile_variants(); ++v) cand.push_back(v);
  }
  struct PS { int64_t ntup; int max_w, b0; };
  std::vector<PS> ps;
  int64_t nops = 0;
  for (auto &pg : plan.progs) {
    int fb[kMaxLines + 1];
    for (int i = 0, f = pg.fb_ix; i < m; ++i, f /= 2) fb[i] = f % 2;
    int64_t ntup = 1;
    for (int x = 0; x < m - 1; ++x) ntup *= (int64_t)c->ps.get(fb[x + 1], (pg.cls >> x) & 1).size();
    ps.push_back({ntup, std::max(1, pg.max_w), pg.b0});
    nops += (int64_t)pg.ops.size();
  }
  // every device buffer holds at least one element (DevBuf), as get_plan allocates them
  auto n1 = [](int64_t n) { return std::max<int64_t>(n, 1); };
  int64_t bytes = 8 * n1(plan.ftab_size) + n1(nops) * (int64_t)sizeof(DOp) + n1((int64_t)tab_jobs(plan.tables, c).size()) * (int64_t)sizeof(DTabJob) +
                  (int64_t)sizeof(int);  // the (empty) tuple list
  for (int v : cand) {
    const int NT = tile_columns(v);
    int64_t est_tiles = 0;
    bool ok = true;
    for (auto &q : ps) {
      if (q.max_w > NT) { ok = false; break; }
      const int per = std::max(1, NT / q.max_w);
      est_tiles += (q.ntup + per - 1) / per;
    }
    if (!ok) continue;
    const int ct = chunk_tiles(est_tiles);
    int64_t chunks = 0, partial = 0;
    for (auto &q : ps) {
      const int per = std::max(1, NT / q.max_w);
      const int64_t nt = (q.ntup + per - 1) / per, nch = (nt + ct - 1) / ct;
      chunks += nch;
      partial += nch * (int64_t)P.r * q.b0 * q.b0;
    }
    bytes += n1(chunks) * (int64_t)(sizeof(DChunk) + sizeof(int64_t) + sizeof(int)) + 8 * n1(partial) +
             n1((int64_t)ps.size()) * (int64_t)sizeof(DProg) + n1(P.nb) * (int64_t)sizeof(I2);
  }
  return bytes;
}

// Device bytes a context holds now: resident tables, Dyson state and its plans (every variant
// still allocated).
int64_t ctx_device_bytes(const xca_ctx *c) {
  int64_t b = ctx_bytes(c) + (c->dyson ? dyson_bytes(c) : 0);
  for (auto &kv : c->plans) b += kv.second->device_bytes;
  for (auto &kv : c->wide_plans) b += kv.second->device_bytes;
  if (c->widectx) b += 8 * (int64_t)c->widectx->M2.n;
  return b;
}

}  // namespace xca

extern "C" xca_ctx *xca_ctx_create(const xca_problem *desc, int device) {
  xca_ctx *c = nullptr;
  try {
    c = ctx_init_host(desc);
    ctx_init_device(c, device);
    return c;
  } catch (std::exception &e) {
    xca::set_error(e.what());
    xca_ctx_destroy(c);
    return nullptr;
  }
}

extern "C" void xca_ctx_destroy(xca_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->plans.clear();
  c->corr_plans.clear();
  c->sigall.reset();
  c->wide_plans.clear();
  c->widectx.reset();
  xca_dyson_release(c);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" int64_t xca_num_backbones(xca_ctx *c, int m) { return c ? c->M.num_backbones(m) : -1; }
extern "C" int64_t xca_num_cached_plans(xca_ctx *c) { return c ? (int64_t)c->plans.size() : -1; }
extern "C" int xca_num_sym_sets(xca_ctx *c) { return c ? c->M.S.q : -1; }

extern "C" int xca_set_G(xca_ctx *c, const double *G) {
  XCA_TRY({
    HIP_CHECK(hipSetDevice(c->device));
    const Problem &P = c->M.P;
    c->G_zero.assign(P.nb, 0);
    for (int b = 0; b < P.nb; ++b) {  // zero blocks prune paths (block_sparse.cpp:45-49)
      double mx = 0;
      const double *g = G + c->M.G_off[b];
      for (int64_t x = 0; x < (int64_t)P.r * P.dims[b] * P.dims[b]; ++x) mx = std::max(mx, std::fabs(g[x]));
      c->G_zero[b] = !(mx >= 1e-16);
    }
    HIP_CHECK(hipMemcpyAsync(c->G.ptr, G, c->M.G_size * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    c->have_G = true;
    return 0;
  })
}

extern "C" int xca_sigma_zero(xca_ctx *c) {
  XCA_TRY({
    HIP_CHECK(launch_fill(c->M.G_size, 0.0, c->sigma.ptr, c->stream));
    return 0;
  })
}

extern "C" int xca_sigma_accumulate(xca_ctx *c, int m, const int *topo, double scale) {
  XCA_TRY({
    accumulate(c, m, topo, nullptr, 0, scale);
    return 0;
  })
}

extern "C" int xca_get_sigma(xca_ctx *c, double *out) {
  XCA_TRY({
    HIP_CHECK(hipMemcpyAsync(out, c->sigma.ptr, c->M.G_size * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    // zero Sigma blocks of zero G blocks are returned as zeros (block_sparse_backbone.cpp:388-395)
    return 0;
  })
}

extern "C" int xca_synchronize(xca_ctx *c) {
  XCA_TRY({
    HIP_CHECK(hipStreamSynchronize(c->stream));
    float ms = 0;
    if (c->last_items > 0 && hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) c->last_ms = ms;
    return 0;
  })
}

extern "C" int xca_sigma(xca_ctx *c, const double *G, int m, const int *topo, const int32_t *fix, int64_t nfix, double *out) {
  XCA_TRY({
    if (xca_set_G(c, G) < 0) return -1;
    HIP_CHECK(launch_fill(c->M.G_size, 0.0, c->sigma.ptr, c->stream));
    accumulate(c, m, topo, fix, fix ? nfix : 0, 1.0);
    if (xca_get_sigma(c, out) < 0) return -1;
    // compute_self_energy returns the blocks whose max |.| is below 1e-16 as zero blocks
    // (set_zero_block_indices, block_sparse.cpp:92-101, called at block_sparse_backbone.cpp:325,336)
    const Problem &P = c->M.P;
    for (int b = 0; b < P.nb; ++b) {
      double *s = out + c->M.G_off[b];
      const int64_t sz = (int64_t)P.r * P.dims[b] * P.dims[b];
      double mx = 0.0;
      for (int64_t x = 0; x < sz; ++x) mx = std::max(mx, std::fabs(s[x]));
      if (mx < 1e-16) std::fill(s, s + sz, 0.0);
    }
    return 0;
  })
}

extern "C" int xca_set_tile_variant(xca_ctx *c, int variant) {
  if (!c) return -1;
  if (variant < -1 || variant >= num_tile_variants()) {
    xca::set_error("xca_set_tile_variant: variant out of range");
    return -1;
  }
  if (hipSetDevice(c->device) != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
    xca::set_error("xca_set_tile_variant: device synchronization failed");
    return -1;
  }
  c->tile_variant = variant;
  c->sigall.reset();
  c->plans.clear();
  c->corr_plans.clear();
  return 0;
}

extern "C" int xca_set_launch_mode(xca_ctx *c, int mode) {
  if (!c) return -1;
  if (mode < 0 || mode > 1) {
    xca::set_error("xca_set_launch_mode: mode must be 0 or 1");
    return -1;
  }
  if (mode != c->launch_mode) {  // explicit-subset plans are laid out for the launch mode
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto e = c->plans.begin(); e != c->plans.end();)
      e = std::find(e->first.begin(), e->first.end(), -1) != e->first.end() ? c->plans.erase(e) : std::next(e);
  }
  c->launch_mode = mode;
  return 0;
}

extern "C" int xca_last_tile(xca_ctx *c, int *variant, int *columns) {
  if (!c) return -1;
  if (variant) *variant = c->last_variant;
  if (columns) *columns = c->last_variant >= 0 ? tile_columns(c->last_variant) : 0;
  return 0;
}

extern "C" int xca_last_stats(xca_ctx *c, int64_t *items, double *macs, double *kernel_ms) {
  XCA_TRY({
    if (items) *items = c->last_items;
    if (macs) *macs = c->last_macs;
    if (kernel_ms) *kernel_ms = c->last_ms;
    return 0;
  })
}

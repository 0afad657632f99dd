// Host orchestration of the device path and the C ABI of the evaluator context.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
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

namespace xca::dev {
int tile_columns();
hipError_t launch_tab(int RB, const DlrFrag &D, const DTabJob *jobs, int njobs, const double *G, const double *kp, const double *km, int p,
                      double *ftab, hipStream_t s);
hipError_t launch_chain(int RB, const void *args, int nchunks, hipStream_t s);
hipError_t launch_fill(int64_t n, double v, double *p, hipStream_t s);
hipError_t launch_reduce(const int64_t *slot, const int *red_chunks, const I2 *red_range, const int64_t *sig_off, const int *sig_size,
                         int nblocks, int max_size, const double *partial, double *sigma, double scale, hipStream_t s);
size_t chain_args_size();

}  // namespace xca::dev

namespace xca {

using namespace xca::dev;

// tiles per chunk: enough chunks to cover the GPU several times, at most 64 tiles per chunk
static int chunk_tiles(int64_t total_tiles) { return (int)std::max<int64_t>(1, std::min<int64_t>(64, total_tiles / 2048)); }

struct DevPlan {
  DevBuf<DOp> ops;
  DevBuf<DProg> progs;
  DevBuf<DTile> tiles;
  DevBuf<I2> chunks;
  DevBuf<int64_t> slot;
  DevBuf<DTabJob> tabjobs;
  DevBuf<int> red_chunks;
  DevBuf<I2> red_range;
  DevBuf<double> ftab, partial;
  int nchunks = 0, ntabjobs = 0;
  int64_t items = 0;
  double macs = 0;
  // host copies kept for the pinned upload lifetime
  std::vector<DOp> h_ops;
  std::vector<DProg> h_progs;
  std::vector<DTile> h_tiles;
  std::vector<I2> h_chunks, h_red_range;
  std::vector<int64_t> h_slot;
  std::vector<DTabJob> h_tab;
  std::vector<int> h_red;
};

}  // namespace xca

using namespace xca;


namespace {

Problem problem_from(const xca_problem *d) {
  if (!d) throw std::invalid_argument("xca_ctx_create: null problem");
  if (d->r <= 0 || d->p <= 0 || d->n <= 0 || d->nblocks <= 0) throw std::invalid_argument("xca_ctx_create: empty problem");
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
  std::vector<double> fLiP, fW, fM2, fAU, fD;
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
  c->frag.upload(all, c->stream);
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

// canonical multiply-add count of one (path, pole tuple) evaluation
double canonical_macs(const Program &pg, int r) {
  double s = 0;
  for (const auto &op : pg.ops) {
    if (op.type == OP_V) s += (double)r * op.bout * op.bin * op.cols;
    else if (op.type == OP_E) s += 4.0 * r * r * op.bin * op.cols + 3.0 * r * op.bin * op.bin * op.cols;
    else s += (double)op.qs * r * op.bin * op.b1 * op.cols + (double)op.qs * op.qs * r * op.bin * op.cols + (double)op.qs * r * op.bout * op.bin * op.cols;
  }
  return s;
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
  auto it = c->plans.find(key);
  if (it != c->plans.end()) return it->second.get();

  const Problem &P = c->M.P;
  const int r = P.r, p = P.p, n = P.n;
  if (m < 1 || m > kMaxLines + 1) throw std::invalid_argument("xca: diagram order must be between 1 and 4");
  Plan plan;
  build_programs(c->M, c->ps, topo, m, 0, c->G_zero, plan);
  auto dp = std::make_unique<DevPlan>();
  const int NT = tile_columns();
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
  // total tile estimate for the chunk size
  int64_t est_tiles = 0;
  for (auto &pg : plan.progs) {
    std::vector<int> fb(m);
    for (int i = 0, f = pg.fb_ix; i < m; ++i, f /= 2) fb[i] = f % 2;
    int64_t nt = 1;
    for (int x = 0; x < nl; ++x) nt *= (int64_t)c->ps.get(fb[x + 1], (pg.cls >> x) & 1).size();
    int per = std::max(1, std::min(NT, NT / std::max(1, pg.max_w)));
    est_tiles += (nt + per - 1) / per;
  }
  for (auto &pg : plan.progs) {
    for (auto &op : pg.ops) {
      if (op.bin > 12 || op.bout > 12 || (op.type == OP_Z && op.qs > kMaxSet))
        throw std::invalid_argument("xca: device path supports blocks of dimension <= 12 and symmetry sets of size <= 8");
    }
    if (pg.max_w > NT) throw std::invalid_argument("xca: block path too wide for one tile");
    DProg d{};
    d.op0 = (int)dp->h_ops.size();
    d.nops = (int)pg.ops.size();
    d.sigma_block = pg.sigma_block;
    d.b0 = pg.b0;
    d.b1 = pg.b1;
    d.g_init = c->M.G_off[pg.init_block];
    d.sign = pg.sign;
    d.nlines = nl;
    std::vector<int> fb(m);
    for (int i = 0, f = pg.fb_ix; i < m; ++i, f /= 2) fb[i] = f % 2;
    int64_t ntup = 1;
    for (int x = 0; x < nl; ++x) {
      d.pk_sign[x] = pg.pk_sign[x];
      d.pk_exp[x] = pg.pk_exp[x];
      int fwd = fb[x + 1], rb = (pg.cls >> x) & 1;
      d.set_off[x] = c->plist_off[(fwd ? 0 : 2) + rb];
      d.set_len[x] = (int)c->ps.get(fwd, rb).size();
      ntup *= d.set_len[x];
    }
    // explicit subset: keep the requested tuples of this program's class, as a plist of tuple
    // indices in the class enumeration
    std::vector<int> tuples;
    if (fix) {
      auto w = wanted.find({pg.fb_ix, pg.o_ix});
      if (w == wanted.end()) continue;
      for (int pix : w->second) {
        int u = 0, stride = 1, cls = 0;
        int pp = pix;
        bool ok = true;
        for (int x = 0; x < nl; ++x) {
          int l = pp % p;
          pp /= p;
          int fwd = fb[x + 1];
          double wl = P.poles[l] * P.beta;
          int rb = fwd ? (wl <= 0 ? 0 : 1) : (wl >= 0 ? 0 : 1);
          cls |= rb << x;
          const auto &lst = c->ps.get(fwd, rb);
          int pos = (int)(std::find(lst.begin(), lst.end(), l) - lst.begin());
          u += pos * stride;
          stride *= (int)lst.size();
        }
        if (cls != pg.cls) ok = false;
        if (ok) tuples.push_back(u);
      }
      if (tuples.empty()) continue;
      ntup = (int64_t)tuples.size();
    }
    // ops
    for (auto &op : pg.ops) {
      DOp o{};
      o.type = op.type; o.bin = op.bin; o.bout = op.bout; o.cols = op.cols;
      o.fbase = op.fbase; o.fstride_l = op.fstride_l;
      o.bar_line = op.bar_line; o.kx = op.kx; o.ks = op.ks; o.nlines = op.nlines;
      for (int q = 0; q < op.nlines; ++q) o.lines[q] = op.lines[q];
      if (op.type == OP_E) {
        const auto &tb = plan.tables[op.etab];
        int64_t esz = 2 * (int64_t)r * P.dims[tb.block] * P.dims[tb.block];
        o.ebase = tb.offset;
        for (int q = 0; q < op.nlines; ++q) o.estride[q] = ipow(p, q) * esz;
      }
      o.mubase = op.mubase; o.qs = op.qs; o.forward = op.forward; o.b1 = op.b1; o.set = op.set;
      if (op.type == OP_Z)
        for (int k = 0; k < op.qs; ++k) o.orb[k] = c->M.S.to_orb[op.set][k];
      dp->h_ops.push_back(o);
    }
    d.max_w = pg.max_w;
    d.items_per_tile = std::max(1, std::min(NT, NT / pg.max_w));
    int pidx = (int)dp->h_progs.size();
    dp->h_progs.push_back(d);
    double macs1 = canonical_macs(pg, r);
    // tiles and chunks
    int64_t per = d.items_per_tile;
    if (!fix) {
      int64_t ntiles = (ntup + per - 1) / per;
      const int ct = chunk_tiles(est_tiles);
      for (int64_t t0 = 0; t0 < ntiles; t0 += ct) {
        I2 ch;
        ch.x = (int)dp->h_tiles.size();
        for (int64_t t = t0; t < std::min(ntiles, t0 + ct); ++t) {
          DTile tl{pidx, (int)(t * per), (int)std::min<int64_t>(per, ntup - t * per), (int)dp->h_chunks.size()};
          dp->h_tiles.push_back(tl);
        }
        ch.y = (int)dp->h_tiles.size();
        dp->h_chunks.push_back(ch);
        dp->h_slot.push_back(0);
      }
    } else {
      // each requested tuple becomes a one-item tile of a one-tuple program copy
      dp->h_progs.pop_back();
      for (int u : tuples) {
        DProg d1 = d;
        d1.set_off[0] = 0;  // patched below
        int decode = u;
        // a program whose pole lists have length 1 selects exactly this tuple
        for (int x = 0; x < nl; ++x) {
          int len = d.set_len[x];
          int pos = decode % len;
          decode /= len;
          d1.set_off[x] = (c->plist_off[(fb[x + 1] ? 0 : 2) + ((pg.cls >> x) & 1)]) + pos;
          d1.set_len[x] = 1;
        }
        int pi = (int)dp->h_progs.size();
        dp->h_progs.push_back(d1);
        I2 ch;
        ch.x = (int)dp->h_tiles.size();
        dp->h_tiles.push_back(DTile{pi, 0, 1, (int)dp->h_chunks.size()});
        ch.y = ch.x + 1;
        dp->h_chunks.push_back(ch);
        dp->h_slot.push_back(0);
      }
    }
    dp->items += ntup;
    dp->macs += macs1 * (double)ntup;
  }
  // partial slots
  int64_t off = 0;
  std::vector<std::vector<int>> per_block(P.nb);
  for (size_t ci = 0; ci < dp->h_chunks.size(); ++ci) {
    const DProg &d = dp->h_progs[dp->h_tiles[dp->h_chunks[ci].x].prog];
    dp->h_slot[ci] = off;
    off += (int64_t)r * d.b0 * d.b0;
    per_block[d.sigma_block].push_back((int)ci);
  }
  for (int b = 0; b < P.nb; ++b) {
    I2 rg;
    rg.x = (int)dp->h_red.size();
    dp->h_red.insert(dp->h_red.end(), per_block[b].begin(), per_block[b].end());
    rg.y = (int)dp->h_red.size();
    dp->h_red_range.push_back(rg);
  }
  // edge-table jobs
  for (const auto &tb : plan.tables) {
    int b = P.dims[tb.block], bb = b * b;
    int per = std::max(1, NT / bb);
    for (int64_t e0 = 0; e0 < tb.nentries; e0 += per) {
      DTabJob j{};
      j.block = tb.block; j.b = b; j.nlines = tb.nlines; j.e0 = (int)e0;
      j.ne = (int)std::min<int64_t>(per, tb.nentries - e0);
      j.g_off = c->M.G_off[tb.block];
      for (int q = 0; q < tb.nlines; ++q) j.lsign[q] = tb.lsign[q];
      j.out_off = tb.offset;
      dp->h_tab.push_back(j);
    }
  }
  hipStream_t s = c->stream;
  dp->ops.upload(dp->h_ops, s);
  dp->progs.upload(dp->h_progs, s);
  dp->tiles.upload(dp->h_tiles, s);
  dp->chunks.upload(dp->h_chunks, s);
  dp->slot.upload(dp->h_slot, s);
  dp->tabjobs.upload(dp->h_tab, s);
  dp->red_chunks.upload(dp->h_red, s);
  dp->red_range.upload(dp->h_red_range, s);
  dp->ftab.alloc(std::max<int64_t>(plan.ftab_size, 1));
  dp->partial.alloc(std::max<int64_t>(off, 1));
  dp->nchunks = (int)dp->h_chunks.size();
  dp->ntabjobs = (int)dp->h_tab.size();
  HIP_CHECK(hipStreamSynchronize(s));
  DevPlan *raw = dp.get();
  c->plans[key] = std::move(dp);
  return raw;
}

void accumulate(xca_ctx *c, int m, const int *topo, const int32_t *fix, int64_t nfix, double scale) {
  if (!c->have_G) throw std::runtime_error("xca: G has not been set");
  DevPlan *dp = get_plan(c, m, topo, fix, nfix);
  hipStream_t s = c->stream;
  c->last_items = dp->items;
  c->last_macs = dp->macs;
  c->last_ms = 0;
  if (dp->nchunks == 0) return;
  if (dp->ntabjobs > 0)
    HIP_CHECK(launch_tab(c->RB, c->D, dp->tabjobs.ptr, dp->ntabjobs, c->G.ptr, c->kp.ptr, c->km.ptr, c->M.P.p, dp->ftab.ptr, s));
  ChainArgs A{};
  A.D = c->D;
  A.ops = dp->ops.ptr; A.progs = dp->progs.ptr; A.tiles = dp->tiles.ptr; A.chunks = dp->chunks.ptr; A.chunk_slot = dp->slot.ptr;
  A.plist = c->plist.ptr; A.G = c->G.ptr; A.opmats = c->opmats.ptr; A.ftab = dp->ftab.ptr; A.kp = c->kp.ptr; A.km = c->km.ptr;
  A.delta = c->delta.ptr; A.delta_refl = c->delta_refl.ptr; A.kpow = c->kpow.ptr; A.partial = dp->partial.ptr;
  A.r = c->M.P.r; A.p = c->M.P.p; A.n = c->M.P.n; A.maxexp = c->maxexp;
  HIP_CHECK(hipEventRecord(c->ev0, s));
  HIP_CHECK(launch_chain(c->RB, &A, dp->nchunks, s));
  HIP_CHECK(hipEventRecord(c->ev1, s));
  int maxsz = 0;
  for (int b = 0; b < c->M.P.nb; ++b) maxsz = std::max(maxsz, c->M.P.r * c->M.P.dims[b] * c->M.P.dims[b]);
  HIP_CHECK(launch_reduce(dp->slot.ptr, dp->red_chunks.ptr, dp->red_range.ptr, c->sig_off.ptr, c->sig_size.ptr, c->M.P.nb, maxsz,
                          dp->partial.ptr, c->sigma.ptr, scale, s));
}

}  // namespace

namespace xca {

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

extern "C" xca_ctx *xca_ctx_create(const xca_problem *desc, int device) {
  xca_ctx *c = nullptr;
  try {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw std::runtime_error("xca_ctx_create: no GPU device available");
    if (device < 0 || device >= ndev) throw std::runtime_error("xca_ctx_create: device index out of range");
    c = new xca_ctx();
    c->device = device;
    HIP_CHECK(hipSetDevice(device));
    c->M.init(problem_from(desc));
    const Problem &P = c->M.P;
    std::vector<double> w(P.p);
    for (int l = 0; l < P.p; ++l) w[l] = P.poles[l] * P.beta;
    c->ps.init(w);
    // padded DLR rank: 32 or 64 rows (16-row MFMA blocks; the row-partitioned GEMMs need the row-block
    // count to divide the 8 waves of a workgroup)
    c->RB = (P.r <= 32) ? 2 : 4;
    if (P.r > 64) throw std::invalid_argument("xca: DLR rank above 64 is not supported by the device path");
    c->RP = 16 * c->RB;
    HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreate(&c->ev0));
    HIP_CHECK(hipEventCreate(&c->ev1));
    hipStream_t s = c->stream;
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
    return c;
  } catch (std::exception &e) {
    xca::set_error(e.what());
    delete c;
    return nullptr;
  }
}

extern "C" void xca_ctx_destroy(xca_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->plans.clear();
  xca_dyson_release(c);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" int64_t xca_num_backbones(xca_ctx *c, int m) { return c ? c->M.num_backbones(m) : -1; }
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
    return xca_get_sigma(c, out);
  })
}

extern "C" int xca_last_stats(xca_ctx *c, int64_t *items, double *macs, double *kernel_ms) {
  XCA_TRY({
    if (items) *items = c->last_items;
    if (macs) *macs = c->last_macs;
    if (kernel_ms) *kernel_ms = c->last_ms;
    return 0;
  })
}

// Host orchestration of the large-block Σ engine. Programs come from the same block-path
// enumeration as the tile engine (build_programs, following block_sparse_backbone.cpp:229-367); each
// program runs over all its pole tuples at once as batched MFMA GEMMs with T resident in HBM:
//   vertex      T[u](t) <- K(t, pole) F T[u](t)                 (batched over items and nodes)
//   edge        g = U^{-1} (L^{-1}P) T, Wg = W g                 (along the node index, wv2c_kernel)
//               c_k = -(Wf_k g_k + f_k Wg_k), d_k = f_k g_k      (batched over items and k)
//               T <- beta [cf2it | tcf2it] [c; d]
//   zero pair   T <- sum_mu F_mu T Fh_mu(t), Fh_mu(t) = sum_kap h(t, mu, kap) F_kap, as two GEMMs:
//               Y = T [Fh_0 | .. | Fh_q-1] and T <- [F_mu columns interleaved] Y
//   final       per-item max |T| (the reference's 1e-16 skip) and Sigma[b0] += scale sum pref T.
// The launch list of a (topology, zero-block pattern, flat-index subset) is built once and replayed.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "host/model.hpp"
#include "plan.hpp"

using namespace xca;

namespace {

int64_t ipow(int64_t a, int e) {
  int64_t r = 1;
  for (int i = 0; i < e; ++i) r *= a;
  return r;
}

void init_wide_ctx(xca_ctx *c) {
  if (c->widectx) return;
  const auto &o = c->M.ops;
  const int r = o.r;
  std::vector<double> M2((size_t)r * 2 * r, 0.0);
  for (int i = 0; i < r; ++i)
    for (int k = 0; k < r; ++k) {
      M2[(size_t)i * 2 * r + k] = c->M.P.beta * o.cf2it[(size_t)i * r + k];
      M2[(size_t)i * 2 * r + r + k] = c->M.P.beta * o.tcf2it[(size_t)i * r + k];
    }
  auto w = std::make_unique<WideCtx>();
  w->M2.upload(M2, c->stream);
  HIP_CHECK(hipStreamSynchronize(c->stream));
  c->widectx = std::move(w);
}

WGemm gemm(const double *A, const double *B, double *C, int M, int N, int K, int lda, int ldb, int ldc, int nitem, int nt) {
  WGemm g{};
  g.A = A; g.B = B; g.C = C;
  g.M = M; g.N = N; g.K = K; g.lda = lda; g.ldb = ldb; g.ldc = ldc;
  g.nitem = nitem; g.nt = nt;
  g.alpha = 1.0; g.beta = 0.0;
  return g;
}

WLaunch of(const WGemm &g) {
  WLaunch L{};
  L.kind = W_GEMM;
  L.g = g;
  return L;
}

// Launch-list builder of the large-block engine, shared by Σ plans and correlator plans: edge
// tables, then programs over their items (pole tuples, first line slowest), each ending either in
// the Σ skip rule and accumulation or in a store of pref * T with the node index fastest. Device
// pointers into the per-item offsets, prefactors and maxima are patched after their upload; every
// operand of every launch is checked against its buffer on the host.
struct WideBuilder {
  xca_ctx *c;
  const std::vector<EdgeTable> &tables;
  double *ftab = nullptr;
  std::vector<WLaunch> launches;
  std::vector<int64_t> offs;
  std::vector<double> prefs;
  int64_t nvmax = 0;
  struct Pending {  // launches that take device pointers into offs / pref / vmax, fixed after upload
    size_t launch;
    int64_t a_off, s_off, pref, vmax, f_off = -1;
  };
  std::vector<Pending> pend;
  std::map<int64_t, int64_t> zperm_of;
  std::vector<double> zperm;
  int64_t max_ws = 0, max_fh = 0, max_yz = 0, max_tab = 0;
  double *T0 = nullptr, *T1 = nullptr, *Y = nullptr, *WG = nullptr, *CD = nullptr, *Fh = nullptr, *Tab = nullptr, *Yz = nullptr;
  size_t zero_launch = (size_t)-1;
  const double *store = nullptr;  // correlator store (output of store-mode programs), checked as a buffer
  int64_t store_n = 0;
  int64_t items = 0;
  double macs = 0;

  WideBuilder(xca_ctx *ctx, const std::vector<EdgeTable> &t) : c(ctx), tables(t) {}

  // first pass: workspace of the tables and of a program over ni items
  void size_tables() {
    const Problem &P = c->M.P;
    for (const auto &tb : tables) max_tab = std::max(max_tab, tb.nentries * P.r * (int64_t)P.dims[tb.block] * P.dims[tb.block]);
  }
  void size_program(const Program &pg, int64_t ni) {
    const int r = c->M.P.r;
    int rows = pg.b1, cols = pg.b1;
    int64_t w = (int64_t)rows * cols;
    for (const Op &op : pg.ops) {
      w = std::max(w, std::max((int64_t)op.bin * cols, (int64_t)op.bout * cols));
      if (op.type == OP_Z) {
        w = std::max(w, (int64_t)op.bout * op.cols);
        max_fh = std::max(max_fh, (int64_t)op.qs * r * op.b1 * op.cols);
        max_yz = std::max(max_yz, ni * r * op.bin * op.qs * op.cols);  // T [Fh_0 | Fh_1 | ...]
        cols = op.cols;
      }
      rows = op.bout;
    }
    max_ws = std::max(max_ws, ni * r * w);
  }
  // the workspace: T0, T1, Y, WG (max_ws each), CD (2 max_ws), Fh, table scratch, zero-pair products
  void alloc_work(WidePlan *wp) {
    const int64_t nwork = 6 * max_ws + max_fh + max_tab + max_yz;
    wp->work.alloc(std::max<int64_t>(nwork, 1));
    T0 = wp->work.ptr;
    T1 = T0 + max_ws;
    Y = T1 + max_ws;
    WG = Y + max_ws;
    CD = WG + max_ws;
    Fh = CD + 2 * max_ws;
    Tab = Fh + max_fh;
    Yz = Tab + max_tab;
  }

  // edge tables: f = vals2coefs(G_b prod K) and W f (the tile engine's tab_kernel), of G_b or its
  // transpose
  void add_tables() {
    const Problem &P = c->M.P;
    const int r = P.r, p = P.p;
    for (const auto &tb : tables) {
      const int b = P.dims[tb.block];
      const int64_t bb = (int64_t)b * b;
      WLaunch L{};
      L.kind = W_GK;
      L.src = c->G.ptr + c->M.G_off[tb.block];
      L.dst = Tab;
      L.n = bb;
      L.ld = tb.tr ? b : 0;
      L.nitem = (int)tb.nentries;
      L.nlines = tb.nlines;
      for (int q = 0; q < 3; ++q) L.ls[q] = q < tb.nlines ? tb.lsign[q] : 0;
      L.p = p;
      launches.push_back(L);
      WLaunch V{};
      V.kind = W_V2C;
      V.v.D = c->D;
      V.v.X = Tab;
      V.v.G = ftab + tb.offset;
      V.v.WG = ftab + tb.offset + r * bb;
      V.v.xsi = r * bb;
      V.v.gsi = V.v.wsi = 2 * r * bb;
      V.v.ld = (int)bb;
      V.v.ncol = (int)bb;
      V.v.nitem = (int)tb.nentries;
      launches.push_back(V);
    }
  }

  // the per-item maxima of Σ programs start at zero
  void add_zero() {
    zero_launch = launches.size();
    WLaunch Z{};
    Z.kind = W_ZERO;
    launches.push_back(Z);
  }

  // one program over its items; out == nullptr: the Σ skip rule and accumulation into the
  // program's Σ block; else pref * T stored at out (item, element, node: node index fastest)
  void add_program(const Program &pg, const std::vector<std::vector<int>> &poles, double *out) {
    const Problem &P = c->M.P;
    const int r = P.r, p = P.p, n = P.n, nl = pg.m - 1;
    const WideCtx &W = *c->widectx;
    const int ni = (int)poles.size();
    // prefactor K(0)^(-e) per line and the sign (hyb.cpp:75-92, block_sparse_backbone.cpp:359-363)
    const int64_t pref_off = (int64_t)prefs.size();
    for (const auto &ls : poles) {
      double pr = (double)pg.sign;
      for (int x = 0; x < nl; ++x)
        if (pg.pk_exp[x] != 0) pr *= std::pow(pg.pk_sign[x] > 0 ? c->M.k0p[ls[x]] : c->M.k0m[ls[x]], -pg.pk_exp[x]);
      prefs.push_back(pr);
    }
    WLaunch I{};
    I.kind = W_INIT;
    I.src = c->G.ptr + c->M.G_off[pg.init_block];
    I.dst = T0;
    I.n = (int64_t)r * pg.b1 * pg.b1;
    I.ld = pg.gtr ? pg.b1 : 0;
    I.nitem = ni;
    launches.push_back(I);
    double *cur = T0, *oth = T1;
    int rows = pg.b1, cols = pg.b1;
    for (const Op &op : pg.ops) {
      if (op.type == OP_V) {
        WGemm g = gemm(c->opmats.ptr + op.fbase, cur, oth, op.bout, cols, op.bin, op.bin, cols, cols, ni, r);
        g.b_si = (int64_t)r * op.bin * cols;
        g.b_st = (int64_t)op.bin * cols;
        g.c_si = (int64_t)r * op.bout * cols;
        g.c_st = (int64_t)op.bout * cols;
        Pending pd{launches.size(), -1, -1, -1, -1};
        if (op.bar_line >= 0) {
          pd.a_off = (int64_t)offs.size();
          for (const auto &ls : poles) offs.push_back((int64_t)ls[op.bar_line] * op.fstride_l);
        }
        if (op.ks != 0) {
          g.S = op.ks > 0 ? c->kp.ptr : c->km.ptr;
          pd.s_off = (int64_t)offs.size();
          for (const auto &ls : poles) offs.push_back((int64_t)ls[op.kx] * r);
        }
        launches.push_back(of(g));
        if (pd.a_off >= 0 || pd.s_off >= 0) pend.push_back(pd);
        std::swap(cur, oth);
      } else if (op.type == OP_E) {
        const int b = op.bin;
        const int64_t w = (int64_t)b * cols;
        WLaunch V{};
        V.kind = W_V2C;
        V.v.D = c->D;
        V.v.X = cur;
        V.v.G = Y;
        V.v.WG = WG;
        V.v.xsi = V.v.gsi = V.v.wsi = r * w;
        V.v.ld = (int)w;
        V.v.ncol = (int)w;
        V.v.nitem = ni;
        launches.push_back(V);
        // table entry (f part) of every item
        const int64_t eoff = (int64_t)offs.size();
        const auto &tb = tables[op.etab];
        const int64_t esz = 2 * (int64_t)r * b * b;
        for (const auto &ls : poles) {
          int64_t e = tb.offset;
          for (int q = 0; q < op.nlines; ++q) e += (int64_t)ls[op.lines[q]] * ipow(p, q) * esz;
          offs.push_back(e);
        }
        // c = -(Wf g + f Wg), d = f g, per item and coefficient index k
        WLaunch Lc{};
        Lc.kind = W_LOCAL;
        Lc.loc.F = ftab;
        Lc.loc.Gk = Y;
        Lc.loc.WGk = WG;
        Lc.loc.C = CD;
        Lc.loc.D = CD + r * w;
        Lc.loc.g_si = r * w;
        Lc.loc.c_si = 2 * r * w;
        Lc.loc.wf_shift = (int64_t)r * b * b;
        Lc.loc.b = b;
        Lc.loc.cols = cols;
        Lc.loc.r = r;
        Lc.loc.nitem = ni;
        Pending pl{launches.size(), -1, -1, -1, -1};
        pl.f_off = eoff;
        pend.push_back(pl);
        launches.push_back(Lc);
        WGemm h = gemm(W.M2.ptr, CD, oth, r, (int)w, 2 * r, 2 * r, (int)w, (int)w, ni, 1);
        h.b_si = 2 * r * w;
        h.c_si = r * w;
        launches.push_back(of(h));
        std::swap(cur, oth);
      } else {  // zero-vertex pair (block_sparse_backbone.cpp:187-227)
        const int bv = op.bin, bvp = op.bout, b1 = op.b1, b0 = op.cols, qs = op.qs;
        if (!zperm_of.count(op.mubase)) {
          // mu operators with their columns reordered to (a qs + mu), so that the sum over mu is one
          // GEMM against T [Fh_0 | Fh_1 | ...] read as (b_v qs) x b0
          zperm_of[op.mubase] = (int64_t)zperm.size();
          const size_t base = zperm.size();
          zperm.resize(base + (size_t)bvp * bv * qs);
          for (int mu = 0; mu < qs; ++mu)
            for (int i = 0; i < bvp; ++i)
              for (int a = 0; a < bv; ++a)
                zperm[base + (size_t)i * bv * qs + (size_t)a * qs + mu] = c->M.opmats[op.mubase + ((size_t)mu * bvp + i) * bv + a];
        }
        WLaunch F{};
        F.kind = W_FH;
        F.dst = Fh;
        F.src = c->opmats.ptr + op.fbase;
        F.qs = qs;
        F.b1 = b1;
        F.b0 = b0;
        F.forward = op.forward;
        F.nflav = n;
        for (int k = 0; k < qs; ++k) F.orb[k] = c->M.S.to_orb[op.set][k];
        launches.push_back(F);
        const int qb0 = qs * b0;
        WGemm g1 = gemm(cur, Fh, Yz, bv, qb0, b1, b1, qb0, qb0, ni, r);  // T(t) [Fh_0(t) | ... ]
        g1.a_si = (int64_t)r * bv * b1;
        g1.a_st = (int64_t)bv * b1;
        g1.b_st = (int64_t)b1 * qb0;
        g1.c_si = (int64_t)r * bv * qb0;
        g1.c_st = (int64_t)bv * qb0;
        launches.push_back(of(g1));
        // the permuted mu operators' address is patched when zperm is uploaded (zp_off)
        WGemm g2 = gemm(nullptr, Yz, oth, bvp, b0, bv * qs, bv * qs, b0, b0, ni, r);
        g2.b_si = (int64_t)r * bv * qb0;
        g2.b_st = (int64_t)bv * qb0;
        g2.c_si = (int64_t)r * bvp * b0;
        g2.c_st = (int64_t)bvp * b0;
        zp_fix.push_back({launches.size(), zperm_of[op.mubase]});
        launches.push_back(of(g2));
        cols = b0;
        std::swap(cur, oth);
      }
      rows = op.bout;
    }
    if (!out) {  // final: skip rule and accumulation into the Sigma block
      const int64_t nout = (int64_t)r * pg.b0 * pg.b0;
      WLaunch X{};
      X.kind = W_MAX;
      X.src = cur;
      X.n = nout;
      X.nitem = ni;
      pend.push_back(Pending{launches.size(), -1, -1, -1, nvmax});
      launches.push_back(X);
      WLaunch Fn{};
      Fn.kind = W_FINAL;
      Fn.src = cur;
      Fn.n = nout;
      Fn.nitem = ni;
      Fn.dst = c->sigma.ptr + c->M.G_off[pg.sigma_block];
      pend.push_back(Pending{launches.size(), -1, -1, pref_off, nvmax});
      launches.push_back(Fn);
      nvmax += ni;
    } else {  // store mode: out[(u w + e) r + t] = pref[u] T[u][t][e]
      if (rows != pg.out_rows || cols != pg.out_cols) throw std::logic_error("wide plan: program output shape");
      WLaunch St{};
      St.kind = W_STORE;
      St.src = cur;
      St.dst = out;
      St.n = (int64_t)pg.out_rows * pg.out_cols;
      St.nitem = ni;
      pend.push_back(Pending{launches.size(), -1, -1, pref_off, -1});
      launches.push_back(St);
    }
    items += ni;
    macs += canonical_macs(pg, r) * (double)ni;
  }

  // launches of the permuted mu operators, patched after the upload
  std::vector<std::pair<size_t, int64_t>> zp_fix;

  // bounds of every operand of every launch (the same strides and per-item offsets as the
  // kernels), uploads, and the device pointers of the per-item arrays
  void finish(WidePlan *wp) {
    const int r = c->M.P.r;
    const WideCtx &W = *c->widectx;
    hipStream_t s = c->stream;
    wp->zperm.upload(zperm, s);
    for (const auto &z : zp_fix) launches[z.first].g.A = wp->zperm.ptr + z.second;
    struct Ext {
      const double *base;
      int64_t n;
    };
    const Ext ext[] = {{c->G.ptr, (int64_t)c->G.n}, {c->sigma.ptr, (int64_t)c->sigma.n}, {c->opmats.ptr, (int64_t)c->opmats.n},
                       {c->kp.ptr, (int64_t)c->kp.n}, {c->km.ptr, (int64_t)c->km.n}, {wp->work.ptr, (int64_t)wp->work.n},
                       {ftab, ftab_n}, {W.M2.ptr, (int64_t)W.M2.n}, {wp->zperm.ptr, (int64_t)wp->zperm.n}, {store, store_n}};
    auto in = [&](const double *ptr, int64_t lo, int64_t hi, const char *what) {
      for (const Ext &e : ext)
        if (e.base && ptr >= e.base && ptr < e.base + e.n) {
          if (ptr + lo < e.base || ptr + hi >= e.base + e.n) throw std::logic_error(std::string("wide plan: operand out of range: ") + what);
          return;
        }
      throw std::logic_error(std::string("wide plan: operand outside every buffer: ") + what);
    };
    std::vector<std::pair<int64_t, int64_t>> aoff(launches.size(), {0, 0}), soff(launches.size(), {-1, -1});
    for (const Pending &pd : pend) {
      const WLaunch &L = launches[pd.launch];
      const int ni = L.kind == W_GEMM ? L.g.nitem : (L.kind == W_LOCAL ? L.loc.nitem : L.nitem);
      if (pd.f_off >= 0) {
        auto mm = std::minmax_element(offs.begin() + pd.f_off, offs.begin() + pd.f_off + ni);
        aoff[pd.launch] = {*mm.first, *mm.second};
      }
      if (pd.a_off >= 0) {
        auto mm = std::minmax_element(offs.begin() + pd.a_off, offs.begin() + pd.a_off + ni);
        aoff[pd.launch] = {*mm.first, *mm.second};
      }
      if (pd.s_off >= 0) {
        auto mm = std::minmax_element(offs.begin() + pd.s_off, offs.begin() + pd.s_off + ni);
        soff[pd.launch] = {*mm.first, *mm.second};
      }
      if (pd.pref >= 0 && pd.pref + ni > (int64_t)prefs.size()) throw std::logic_error("wide plan: prefactor range");
      if (pd.vmax >= 0 && pd.vmax + ni > nvmax) throw std::logic_error("wide plan: vmax range");
    }
    for (size_t i = 0; i < launches.size(); ++i) {
      const WLaunch &L = launches[i];
      switch (L.kind) {
        case W_GEMM: {
          const WGemm &g = L.g;
          const int64_t nu = g.nitem - 1, nt = g.nt - 1;
          in(g.A, aoff[i].first, aoff[i].second + nu * g.a_si + nt * g.a_st + (int64_t)(g.M - 1) * g.lda + g.K - 1, "A");
          in(g.B, 0, nu * g.b_si + nt * g.b_st + (int64_t)(g.K - 1) * g.ldb + g.N - 1, "B");
          in(g.C, 0, nu * g.c_si + nt * g.c_st + (int64_t)(g.M - 1) * g.ldc + g.N - 1, "C");
          if (g.S) in(g.S, soff[i].first, soff[i].second + nt, "S");
          if (g.K > g.lda || g.N > g.ldb || g.N > g.ldc) throw std::logic_error("wide plan: leading dimension");
          break;
        }
        case W_INIT:
          in(L.src, 0, L.n - 1, "init src");
          in(L.dst, 0, L.n * L.nitem - 1, "init dst");
          if (L.ld > 0 && (int64_t)r * L.ld * L.ld != L.n) throw std::logic_error("wide plan: transposed init shape");
          break;
        case W_V2C: {
          const WV2C &v = L.v;
          const int64_t span = (int64_t)(r - 1) * v.ld + v.ncol - 1;
          in(v.X, 0, (int64_t)(v.nitem - 1) * v.xsi + span, "v2c X");
          in(v.G, 0, (int64_t)(v.nitem - 1) * v.gsi + span, "v2c G");
          in(v.WG, 0, (int64_t)(v.nitem - 1) * v.wsi + span, "v2c WG");
          if (v.ncol > v.ld) throw std::logic_error("wide plan: leading dimension");
          break;
        }
        case W_LOCAL: {
          const WLocal &q = L.loc;
          const int64_t bb = (int64_t)q.b * q.b, bc = (int64_t)q.b * q.cols;
          in(q.F, aoff[i].first, aoff[i].second + (int64_t)(q.r - 1) * bb + q.wf_shift + bb - 1, "local f");
          in(q.Gk, 0, (int64_t)(q.nitem - 1) * q.g_si + (int64_t)q.r * bc - 1, "local g");
          in(q.WGk, 0, (int64_t)(q.nitem - 1) * q.g_si + (int64_t)q.r * bc - 1, "local Wg");
          in(q.C, 0, (int64_t)(q.nitem - 1) * q.c_si + (int64_t)q.r * bc - 1, "local c");
          in(q.D, 0, (int64_t)(q.nitem - 1) * q.c_si + (int64_t)q.r * bc - 1, "local d");
          if (q.r != r) throw std::logic_error("wide plan: node count");
          break;
        }
        case W_ZERO:
          break;
        case W_GK:
          in(L.src, 0, (int64_t)r * L.n - 1, "gk src");
          in(L.dst, 0, (int64_t)r * L.n * L.nitem - 1, "gk dst");
          if (L.ld > 0 && (int64_t)L.ld * L.ld != L.n) throw std::logic_error("wide plan: transposed table shape");
          break;
        case W_FH:
          in(L.src, 0, (int64_t)L.qs * L.b1 * L.b0 - 1, "fh src");
          in(L.dst, 0, (int64_t)L.qs * r * L.b1 * L.b0 - 1, "fh dst");
          if (L.qs > 8) throw std::logic_error("wide plan: symmetry set above 8");
          break;
        case W_MAX:
          in(L.src, 0, L.n * L.nitem - 1, "max src");
          break;
        case W_FINAL:
          in(L.src, 0, L.n * L.nitem - 1, "final src");
          in(L.dst, 0, L.n - 1, "final dst");
          break;
        case W_STORE:
          in(L.src, 0, (int64_t)r * L.n * L.nitem - 1, "store src");
          in(L.dst, 0, (int64_t)r * L.n * L.nitem - 1, "store dst");
          break;
      }
    }
    wp->offs.upload(offs, s);
    wp->pref.upload(prefs, s);
    wp->vmax.alloc(std::max<int64_t>(nvmax, 1));
    if (zero_launch != (size_t)-1) {
      launches[zero_launch].dst = wp->vmax.ptr;
      launches[zero_launch].n = std::max<int64_t>(nvmax, 1);
    }
    for (const Pending &pd : pend) {
      WLaunch &L = launches[pd.launch];
      if (pd.f_off >= 0) L.loc.f_off = wp->offs.ptr + pd.f_off;
      if (pd.a_off >= 0) L.g.a_off = wp->offs.ptr + pd.a_off;
      if (pd.s_off >= 0) L.g.s_off = wp->offs.ptr + pd.s_off;
      if (pd.pref >= 0) L.pref = wp->pref.ptr + pd.pref;
      if (pd.vmax >= 0) L.vmax = wp->vmax.ptr + pd.vmax;
    }
    wp->launches = std::move(launches);
    wp->items = items;
    wp->macs = macs;
    wp->device_bytes = 8 * (int64_t)(wp->work.n + wp->ftab.n + wp->pref.n + wp->vmax.n + wp->offs.n + wp->zperm.n);
    HIP_CHECK(hipStreamSynchronize(s));
  }
  int64_t ftab_n = 0;
};

// the pole tuples of a program's class, first line slowest (the chain kernel's enumeration)
std::vector<std::vector<int>> class_tuples(const xca_ctx *c, const Program &pg) {
  const int nl = pg.m - 1;
  int fb[kMaxLines + 1];
  for (int i = 0, f = pg.fb_ix; i < pg.m; ++i, f /= 2) fb[i] = f % 2;
  std::vector<const std::vector<int> *> lists(nl);
  for (int x = 0; x < nl; ++x) lists[x] = &c->ps.get(fb[x + 1], (pg.cls >> x) & 1);
  int64_t ntup = 1;
  for (int x = 0; x < nl; ++x) ntup *= (int64_t)lists[x]->size();
  std::vector<std::vector<int>> out;
  out.reserve((size_t)ntup);
  for (int64_t u = 0; u < ntup; ++u) {
    std::vector<int> ls(nl);
    int64_t v = u;
    for (int x = nl - 1; x >= 0; --x) {
      ls[x] = (*lists[x])[v % lists[x]->size()];
      v /= (int64_t)lists[x]->size();
    }
    out.push_back(ls);
  }
  return out;
}

}  // namespace

namespace xca {

static WidePlan *get_wide_plan(xca_ctx *c, int m, const int *topo, const int32_t *fix, int64_t nfix) {
  std::vector<int> key;
  key.push_back(m);
  key.insert(key.end(), topo, topo + 2 * m);
  key.insert(key.end(), c->G_zero.begin(), c->G_zero.end());
  if (fix) {
    key.push_back(-1);
    key.insert(key.end(), fix, fix + nfix);
  }
  auto found = c->wide_plans.find(key);
  if (found != c->wide_plans.end()) return found->second.get();
  if (fix)  // one explicit-subset plan at a time
    for (auto e = c->wide_plans.begin(); e != c->wide_plans.end();) {
      if (std::find(e->first.begin(), e->first.end(), -1) == e->first.end()) {
        ++e;
        continue;
      }
      HIP_CHECK(hipStreamSynchronize(c->stream));  // its launches may still be in flight
      e = c->wide_plans.erase(e);
    }
  if (m < 1 || m > kMaxLines + 1) throw std::invalid_argument("xca: diagram order must be between 1 and 4");
  init_wide_ctx(c);
  const Problem &P = c->M.P;
  const int p = P.p, n = P.n, nl = m - 1;
  Plan plan;
  build_programs(c->M, c->ps, topo, m, 0, c->G_zero, plan);

  // explicit flat-index subsets: (fb_ix, o_ix) -> requested p_ix
  std::map<std::pair<int, int>, std::vector<int>> wanted;
  const int64_t o_max = ipow(n, m - 1), p_max = ipow(p, m - 1), f_max = ipow(2, m) * o_max * p_max;
  if (fix)
    for (int64_t i = 0; i < nfix; ++i) {
      const int64_t f = fix[i];
      if (f < 0 || f >= f_max) throw std::runtime_error("DiagramEvaluator::eval_self_energy: f_ix out of range");
      const int64_t rest = f / o_max;
      wanted[{(int)(rest / p_max), (int)(f % o_max)}].push_back((int)(rest % p_max));
    }

  // items (pole tuples) of every program, first line slowest
  std::vector<std::pair<const Program *, std::vector<std::vector<int>>>> progs;
  for (const Program &pg : plan.progs) {
    std::vector<std::vector<int>> poles;
    if (!fix) {
      poles = class_tuples(c, pg);
    } else {
      int fb[kMaxLines + 1];
      for (int i = 0, f = pg.fb_ix; i < m; ++i, f /= 2) fb[i] = f % 2;
      auto w = wanted.find({pg.fb_ix, pg.o_ix});
      if (w == wanted.end()) continue;
      for (int pix : w->second) {
        std::vector<int> ls(nl);
        int cls = 0, pp = pix;
        for (int x = 0; x < nl; ++x) {
          ls[x] = pp % p;
          pp /= p;
          const double wl = P.poles[ls[x]] * P.beta;
          const int rb = fb[x + 1] ? (wl <= 0 ? 0 : 1) : (wl >= 0 ? 0 : 1);  // backbone.cpp:88,110
          cls |= rb << x;
        }
        if (cls == pg.cls) poles.push_back(ls);
      }
    }
    if (poles.empty()) continue;
    progs.push_back({&pg, std::move(poles)});
  }

  auto wp = std::make_unique<WidePlan>();
  WideBuilder B(c, plan.tables);
  B.size_tables();
  for (const auto &pi : progs) B.size_program(*pi.first, (int64_t)pi.second.size());
  B.alloc_work(wp.get());
  wp->ftab.alloc(std::max<int64_t>(plan.ftab_size, 1));
  B.ftab = wp->ftab.ptr;
  B.ftab_n = (int64_t)wp->ftab.n;
  B.add_tables();
  B.add_zero();
  for (const auto &pi : progs) B.add_program(*pi.first, pi.second, nullptr);
  B.finish(wp.get());
  WidePlan *raw = wp.get();
  c->wide_plans[key] = std::move(wp);
  return raw;
}

// Device bytes of the Sigma plan of one topology at the context's zero-block pattern, with every pole
// tuple of every program (the allocations of get_wide_plan, without building it), plus the
// problem-level matrix the engine keeps once per context.
int64_t wide_plan_bytes_estimate(xca_ctx *c, int m, const int *topo) {
  Plan plan;
  build_programs(c->M, c->ps, topo, m, 0, c->G_zero, plan);
  WideBuilder B(c, plan.tables);
  B.size_tables();
  int64_t items = 0, offs = 0, zperm = 0;
  std::map<int64_t, int> mus;
  for (const Program &pg : plan.progs) {
    int fb[kMaxLines + 1];
    for (int i = 0, f = pg.fb_ix; i < pg.m; ++i, f /= 2) fb[i] = f % 2;
    int64_t ni = 1;
    for (int x = 0; x < pg.m - 1; ++x) ni *= (int64_t)c->ps.get(fb[x + 1], (pg.cls >> x) & 1).size();
    if (ni == 0) continue;
    B.size_program(pg, ni);
    items += ni;
    for (const Op &op : pg.ops) {
      if (op.type == OP_V) offs += ni * ((op.bar_line >= 0 ? 1 : 0) + (op.ks != 0 ? 1 : 0));
      if (op.type == OP_E) offs += ni;
      if (op.type == OP_Z && !mus.count(op.mubase)) {
        mus[op.mubase] = 1;
        zperm += (int64_t)op.bout * op.bin * op.qs;
      }
    }
  }
  auto n1 = [](int64_t n) { return std::max<int64_t>(n, 1); };
  const int64_t work = n1(6 * B.max_ws + B.max_fh + B.max_tab + B.max_yz);
  const int r = c->M.P.r;
  return 8 * (work + n1(plan.ftab_size) + 2 * n1(items) + n1(offs) + n1(zperm)) + 8 * (int64_t)r * 2 * r;
}

std::unique_ptr<WidePlan> build_wide_corr(xca_ctx *c, const std::vector<EdgeTable> &tables, int64_t ftab_size,
                                          const std::vector<WideCorrProg> &progs, int nbatch, double *store, int64_t store_n) {
  init_wide_ctx(c);
  auto wp = std::make_unique<WidePlan>();
  WideBuilder B(c, tables);
  B.size_tables();
  for (const auto &e : progs) B.size_program(*e.pg, (int64_t)e.poles.size());
  B.alloc_work(wp.get());
  wp->ftab.alloc(std::max<int64_t>(ftab_size, 1));
  B.ftab = wp->ftab.ptr;
  B.ftab_n = (int64_t)wp->ftab.n;
  B.store = store;
  B.store_n = store_n;
  B.add_tables();
  wp->seg.push_back(B.launches.size());
  for (int bi = 0; bi < nbatch; ++bi) {
    for (const auto &e : progs)
      if (e.batch == bi && !e.poles.empty()) {
        if (e.out_off < 0 || e.out_off + (int64_t)c->M.P.r * e.pg->out_rows * e.pg->out_cols * (int64_t)e.poles.size() > store_n)
          throw std::logic_error("wide plan: correlator store range");
        B.add_program(*e.pg, e.poles, store + e.out_off);
      }
    wp->seg.push_back(B.launches.size());
  }
  B.finish(wp.get());
  return wp;
}

void wide_launch_range(xca_ctx *c, const WidePlan &wp, size_t l0, size_t l1) {
  WArgs A{c->kp.ptr, c->km.ptr, c->delta.ptr, c->delta_refl.ptr, c->M.P.r, c->RB};
  for (size_t i = l0; i < l1 && i < wp.launches.size(); ++i) HIP_CHECK(launch_wide(wp.launches[i], A, 1.0, c->stream));
}

void wide_accumulate(xca_ctx *c, int m, const int *topo, const int32_t *fix, int64_t nfix, double scale) {
  WidePlan *wp = get_wide_plan(c, m, topo, fix, nfix);
  const WideCtx &W = *c->widectx;
  WArgs A{c->kp.ptr, c->km.ptr, c->delta.ptr, c->delta_refl.ptr, c->M.P.r, c->RB};
  hipStream_t s = c->stream;
  c->last_items = wp->items;
  c->last_macs = wp->macs;
  c->last_ms = 0;
  c->last_variant = -1;
  if (wp->items == 0) return;
  if (!c->outer_events) HIP_CHECK(hipEventRecord(c->ev0, s));
  for (const WLaunch &L : wp->launches) HIP_CHECK(launch_wide(L, A, scale, s));
  if (!c->outer_events) HIP_CHECK(hipEventRecord(c->ev1, s));
}

}  // namespace xca

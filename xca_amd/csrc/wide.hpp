// Large-block Σ engine: block paths whose blocks exceed the tile engine's 12 rows (the 5-orbital d
// shell reaches 100) run as sequences of batched FP64 MFMA GEMMs over the items (pole tuples) of one
// program, with T in HBM. Shared by the kernels (wide.hip) and the host orchestration (wide.cpp).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace xca::dev {

// Batched row-major GEMM: for item u < nitem and t < nt,
//   C[u,t] = alpha * s(u,t) * A[u,t] B[u,t] + beta * C[u,t]
// with X[u,t] = X + x_off[u] (when given) + u x_si + t x_st and s(u,t) = S[s_off[u] + t] (when S is
// given, else 1). A is M x K (lda), B is K x N (ldb), C is M x N (ldc). beta == 0 never reads C.
struct WGemm {
  const double *A, *B;
  double *C;
  const double *S;
  const int64_t *a_off, *s_off;
  int64_t a_si, a_st, b_si, b_st, c_si, c_st;
  int M, N, K, lda, ldb, ldc;
  int nitem, nt;
  double alpha, beta;
};

enum WKind : int { W_GEMM = 0, W_INIT, W_USOLVE, W_GK, W_FH, W_MAX, W_FINAL };

// One launch of a wide plan. Fields by kind:
//   W_INIT:   dst[u] = src (n doubles), optionally transposed (tr, b x b blocks of r nodes)
//   W_USOLVE: X[u] (r rows, ncol columns, row stride ld, item stride si) <- U^{-1} X[u]
//   W_GK:     dst[e] (r x n) = G block src (r x n) times prod_q K^{ls_q}(t, pole of line q of entry e)
//   W_FH:     dst[mu][t] (b1 x b0) = sum_kap h(t, mu, kap) F_kap (F_kap at src + kap b1 b0)
//   W_MAX:    vmax[u] = max |src[u]| over n doubles
//   W_FINAL:  dst (r b0^2, the Sigma block) += scale sum_u [vmax[u] |pref[u]| > 1e-16] pref[u] src[u]
struct WLaunch {
  int kind;
  int nitem, ncol, ld;
  int64_t n, si;
  const double *src;
  double *dst;
  WGemm g;
  // W_GK
  int nlines, ls[3], p;
  // W_FH
  int qs, b1, b0, forward, nflav, orb[8];
  // W_MAX / W_FINAL
  double *vmax;
  const double *pref;
};

struct WArgs {
  const double *U;      // r x r upper factor of cf2it (row-major)
  const double *kp, *km, *delta, *delta_refl;
  int r, pad0;
};

hipError_t launch_wide(const WLaunch &L, const WArgs &A, double scale, hipStream_t s);

}  // namespace xca::dev

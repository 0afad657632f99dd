// Large-block Σ engine: block paths whose blocks exceed the tile engine's 12 rows (the 5-orbital d
// shell reaches 100) run as sequences of batched FP64 MFMA GEMMs over the items (pole tuples) of one
// program, with T in HBM. Shared by the kernels (wide.hip) and the host orchestration (wide.cpp).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>

#include "engine.hpp"

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

// Node-local products of an edge for item u and node index k (f_k, Wf_k: b x b table rows of the
// item's entry at F + f_off[u] + k b^2 and + r b^2; g_k, Wg_k: b x cols):
//   C[u][k] = -(Wf_k g_k + f_k Wg_k),  D[u][k] = f_k g_k.
struct WLocal {
  const double *F, *Gk, *WGk;
  double *C, *D;
  const int64_t *f_off;
  int64_t g_si, c_si;   // item strides of g/Wg and of C/D (node stride = b cols)
  int64_t wf_shift;     // r b^2
  int b, cols, r, nitem;
};

enum WKind : int { W_GEMM = 0, W_INIT, W_V2C, W_GK, W_FH, W_MAX, W_FINAL, W_LOCAL, W_ZERO, W_STORE };

// One launch of a wide plan. Fields by kind:
//   W_INIT:   dst[u] = src (n doubles); ld > 0: src holds r blocks of ld x ld, each copied transposed
//   W_V2C:    v.G = vals2coefs(v.X), v.WG = W v.G (launch_wv2c)
//   W_LOCAL:  the node-local products of an edge
//   W_ZERO:   dst[0, n) = 0
//   W_GK:     dst[e] (r x n) = G block src (r x n) times prod_q K^{ls_q}(t, pole of line q of entry e);
//             ld > 0: the block transposed (n = ld^2)
//   W_FH:     dst[t] (b1 x qs b0), block mu = sum_kap h(t, mu, kap) F_kap (F_kap at src + kap b1 b0)
//   W_MAX:    vmax[u] = max(vmax[u], max |src[u]|) over n doubles (vmax zeroed by W_ZERO first)
//   W_FINAL:  dst (r b0^2, the Sigma block) += scale sum_u [vmax[u] |pref[u]| > 1e-16] pref[u] src[u]
//   W_STORE:  dst[(u n + e) r + t] = pref[u] src[(u r + t) n + e]  (the item's T of n elements per node,
//             node index fastest: the correlator store)
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
  WV2C v;
  WLocal loc;
};

struct WArgs {
  const double *kp, *km, *delta, *delta_refl;
  int r, RB;
};

hipError_t launch_wv2c(int RB, const WV2C &A, hipStream_t s);

hipError_t launch_wide(const WLaunch &L, const WArgs &A, double scale, hipStream_t s);

}  // namespace xca::dev

// Pseudo-particle Dyson equation and the self-consistency update on the device (FP64), batched over
// the blocks of many problems (dyson_it_ppsc::solve, c++/triqs_xca/dlr_dyson_ppsc.hpp:135-168, and
// BlockSparseSolver.solve, python/triqs_xca/block_sparse_solver.py:295-346,427-449,572-590).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "dyson.hpp"

namespace xca::dev {

// segment s of a batch holding global index k: the last s with prefix[s] <= k
__device__ __forceinline__ static int find_seg(const int *prefix, int nsegs, int k) {
  int lo = 0, hi = nsegs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= k) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// value of lane k (k wave-uniform)
__device__ __forceinline__ double lane_value(double v, int k) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), k), hi = __builtin_amdgcn_readlane(__double2hiint(v), k);
  return __hiloint2double(hi, lo);
}

// One step of the wave argmax (largest value, then smallest index) over the DPP pattern CTRL,
// restricted to the rows in ROW_MASK; lanes without a source combine with (-1, idx0), which never wins.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void dpp_argmax(double &v, int &i, int idx0) {
  const double none = -1.0;
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(none), __double2loint(v), CTRL, ROW_MASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(none), __double2hiint(v), CTRL, ROW_MASK, 0xF, false);
  const int i2 = __builtin_amdgcn_update_dpp(idx0, i, CTRL, ROW_MASK, 0xF, false);
  const double v2 = __hiloint2double(hi, lo);
  if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
}

// (largest value, smallest index among equal values) of the wave, in every lane: a reduction to
// lane 63 by row shifts 1, 2, 4, 8 and the row broadcasts of lanes 15 and 31 (no LDS traffic), then
// read from lane 63.
__device__ __forceinline__ void wave_argmax(double &v, int &i, int idx0) {
  dpp_argmax<0x111, 0xF>(v, i, idx0);  // row_shr:1
  dpp_argmax<0x112, 0xF>(v, i, idx0);  // row_shr:2
  dpp_argmax<0x114, 0xF>(v, i, idx0);  // row_shr:4
  dpp_argmax<0x118, 0xF>(v, i, idx0);  // row_shr:8
  dpp_argmax<0x142, 0xA>(v, i, idx0);  // row_bcast:15 into rows 1 and 3
  dpp_argmax<0x143, 0xC>(v, i, idx0);  // row_bcast:31 into rows 2 and 3
  v = lane_value(v, 63);
  i = __builtin_amdgcn_readlane(i, 63);
}

// LU solve of one column by one wave, lane j holding x[j] (j < r <= 64): the row interchanges in
// order, then column-oriented substitutions (the forward pass subtracts its terms in the same order
// as the row-oriented one; the backward pass subtracts them from the last row up). The factors and
// pivots are read from LDS.
__device__ __forceinline__ double lu_solve_wave(const double *__restrict__ lu, const int *__restrict__ piv, int r, double v, int lane) {
  for (int k = 0; k < r; ++k) {
    const int p = __builtin_amdgcn_readfirstlane(piv[k]);
    if (p != k) {
      const double a = lane_value(v, k), b = lane_value(v, p);
      if (lane == k) v = b;
      else if (lane == p) v = a;
    }
  }
  // branch-free steps: every lane loads a factor entry (lanes past r read row 0, inside the r x r
  // factors) so that the loads of later steps can issue early; lanes outside the update keep v
  const int row = lane < r ? lane : 0;
#pragma unroll 4
  for (int k = 0; k + 1 < r; ++k) {
    const double l = lu[row * r + k];
    const double xk = lane_value(v, k);
    const double u = v - l * xk;
    v = (lane > k && lane < r) ? u : v;
  }
#pragma unroll 4
  for (int k = r - 1; k >= 0; --k) {
    const double l = lu[row * r + k], dk = lu[k * r + k];
    if (lane == k) v = v / dk;
    const double xk = lane_value(v, k);
    const double u = v - l * xk;
    v = lane < k ? u : v;
  }
  return v;
}

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma16x16x4(double a, double b, v4d c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

constexpr int kDysonThreads = 512;
constexpr int kDysonMaxNp = 784;  // padded r n: DLR rank <= 64, blocks <= 12
constexpr int kMaxRank = 64;      // DLR rank limit of the device path (checked at context creation)
constexpr int kLdsSysMax = kDysonLdsSysMax;

// Orders the LDS accesses of one wave around this point (a wave executes its LDS operations in
// issue order; this keeps the compiler from moving them across it).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// One 16-column panel of the blocked LU (rows k0..Np-1 of columns k0..k0+15, m = Np - k0 rows of 16
// in LDS), factorised by one wave with partial pivoting: for each column, the first row of maximal
// modulus at or below the diagonal (lane scan, then a DPP reduction) is interchanged with the
// diagonal row, then every row below is updated, its entries read together into registers and
// written back. ipk[jj] receives the absolute pivot row k0 + p (LDS).
__device__ __forceinline__ void panel_factor(double *Pn, int m, int k0, int Np, int *ipk, int lane) {
#pragma unroll
  for (int jj = 0; jj < 16; ++jj) {
    double best = -1.0;
    int bi = Np;
#pragma unroll 4
    for (int i = jj + lane; i < m; i += 64) {
      const double v = fabs(Pn[i * 16 + jj]);
      if (v > best) { best = v; bi = i; }
    }
    wave_argmax(best, bi, Np);
    const int p = bi;
    if (lane == 0) ipk[jj] = k0 + p;
    if (p != jj && lane < 16) {
      const double a = Pn[jj * 16 + lane];
      Pn[jj * 16 + lane] = Pn[p * 16 + lane];
      Pn[p * 16 + lane] = a;
    }
    wave_lds_sync();
    double prow[16];
#pragma unroll
    for (int c = jj; c < 16; ++c) prow[c] = Pn[jj * 16 + c];
    const double pv = prow[jj];
#pragma unroll 2
    for (int i = jj + 1 + lane; i < m; i += 64) {
      double v[16];
#pragma unroll
      for (int c = jj; c < 16; ++c) v[c] = Pn[i * 16 + c];
      const double l = v[jj] / pv;
      v[jj] = l;
#pragma unroll
      for (int c = jj + 1; c < 16; ++c) v[c] -= l * prow[c];
#pragma unroll
      for (int c = jj; c < 16; ++c) Pn[i * 16 + c] = v[c];
    }
    wave_lds_sync();
  }
}

// The same panel factorisation with the rows in registers, for panels of at most 64 kWavePanelSlots
// rows. Rows stay where they were loaded (slot k of a lane holds the row now at position pos[k]); an
// interchange only swaps two positions, and the rows are written to their final positions at the
// end. The pivot is the row of maximal modulus at the smallest position (the first such row of
// panel_factor's scan), found by a DPP reduction over (modulus, position) and broadcast with lane
// reads; the arithmetic is panel_factor's.
constexpr int kWavePanelSlots = 2;  // wave_panel_step's pivot broadcast is written for two slots
static_assert(64 * kWavePanelSlots >= kDysonMultiNp, "one-launch systems must fit the register panel");

template <int JJ>
__device__ __forceinline__ void wave_panel_step(double (&v)[kWavePanelSlots][16], int (&pos)[kWavePanelSlots], int m, int k0, int Np, int *ipk,
                                                int lane) {
  double best = -1.0;
  int bp = Np;
#pragma unroll
  for (int k = 0; k < kWavePanelSlots; ++k)
    if (pos[k] >= JJ && pos[k] < m) {
      const double a = fabs(v[k][JJ]);
      if (a > best || (a == best && pos[k] < bp)) { best = a; bp = pos[k]; }
    }
  wave_argmax(best, bp, Np);
  const int p = bp;
  if (lane == 0) ipk[JJ] = k0 + p;
  const unsigned long long h0 = __ballot(pos[0] == p);
  double prow[16];
  if (h0) {
    const int pl = __builtin_ctzll(h0);
#pragma unroll
    for (int c = JJ; c < 16; ++c) prow[c] = lane_value(v[0][c], pl);
  } else {
    const int pl = __builtin_ctzll(__ballot(pos[1] == p));
#pragma unroll
    for (int c = JJ; c < 16; ++c) prow[c] = lane_value(v[1][c], pl);
  }
  // rows JJ and p interchanged: the row at position JJ moves to p, the pivot row to JJ
#pragma unroll
  for (int k = 0; k < kWavePanelSlots; ++k) pos[k] = pos[k] == p ? JJ : (pos[k] == JJ ? p : pos[k]);
  const double pv = prow[JJ];
#pragma unroll
  for (int k = 0; k < kWavePanelSlots; ++k)
    if (pos[k] > JJ && pos[k] < m) {
      const double l = v[k][JJ] / pv;
      v[k][JJ] = l;
#pragma unroll
      for (int c = JJ + 1; c < 16; ++c) v[k][c] -= l * prow[c];
    }
}

template <int JJ>
__device__ __forceinline__ void wave_panel_steps(double (&v)[kWavePanelSlots][16], int (&pos)[kWavePanelSlots], int m, int k0, int Np, int *ipk,
                                                 int lane) {
  if constexpr (JJ < 16) {
    wave_panel_step<JJ>(v, pos, m, k0, Np, ipk, lane);
    wave_panel_steps<JJ + 1>(v, pos, m, k0, Np, ipk, lane);
  }
}

// Panel of m <= 64 kWavePanelSlots rows at Pn (16 doubles per row) by one wave, through registers.
__device__ __forceinline__ void panel_factor_regs(double *Pn, int m, int k0, int Np, int *ipk, int lane) {
  double v[kWavePanelSlots][16];
  int pos[kWavePanelSlots];
#pragma unroll
  for (int k = 0; k < kWavePanelSlots; ++k) {
    const int i = lane + 64 * k;
    pos[k] = i;
#pragma unroll
    for (int c = 0; c < 16; ++c) v[k][c] = i < m ? Pn[i * 16 + c] : 0.0;
  }
  wave_lds_sync();  // every row is read before any is written back
  wave_panel_steps<0>(v, pos, m, k0, Np, ipk, lane);
#pragma unroll
  for (int k = 0; k < kWavePanelSlots; ++k)
    if (pos[k] < m) {
#pragma unroll
      for (int c = 0; c < 16; ++c) Pn[pos[k] * 16 + c] = v[k][c];
    }
  wave_lds_sync();
}

// The row interchanges of the panel at k0 (absolute rows ipk[0..15], in order) applied to column col
// of the system, and, right of the panel, U12 = L11^{-1} A12 for that column (L11 unit lower, its
// rows 16 doubles apart at L11).
__device__ __forceinline__ void panel_column(double *S, int Np, int k0, int col, const int *ipk, const double *L11) {
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int pr = ipk[q];
    if (pr != k0 + q) {
      const double a = S[(size_t)(k0 + q) * Np + col];
      S[(size_t)(k0 + q) * Np + col] = S[(size_t)pr * Np + col];
      S[(size_t)pr * Np + col] = a;
    }
  }
  if (col >= k0 + 16) {
    double x[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) x[q] = S[(size_t)(k0 + q) * Np + col];
#pragma unroll
    for (int q = 1; q < 16; ++q) {
      double sacc = x[q];
#pragma unroll
      for (int q0 = 0; q0 < 16; ++q0)
        if (q0 < q) sacc -= L11[q * 16 + q0] * x[q0];
      x[q] = sacc;
    }
#pragma unroll
    for (int q = 1; q < 16; ++q) S[(size_t)(k0 + q) * Np + col] = x[q];
  }
}

// A22 -= L21 U12 on the 16 x 16 tile at (ti, tj) in four MFMA steps; row ti + i of L21, panel column
// c, is Lrow[i * ldl + c].
__device__ __forceinline__ void trailing_tile(double *S, int Np, int k0, int ti, int tj, const double *Lrow, int ldl, int lane) {
  v4d acc;
#pragma unroll
  for (int v = 0; v < 4; ++v) acc[v] = S[(size_t)(ti + 4 * v + (lane >> 4)) * Np + tj + (lane & 15)];
#pragma unroll
  for (int ks = 0; ks < 4; ++ks) {
    const double a = -Lrow[(size_t)(lane & 15) * ldl + 4 * ks + (lane >> 4)];
    const double bb = S[(size_t)(k0 + 4 * ks + (lane >> 4)) * Np + tj + (lane & 15)];
    acc = mfma16x16x4(a, bb, acc);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) S[(size_t)(ti + 4 * v + (lane >> 4)) * Np + tj + (lane & 15)] = acc[v];
}

// P1 = A1 B1, and P2 = A2 B2 when TWO, for the r x nc products on 16 x 16 FP64 MFMA tiles (one wave
// per tile at a time, K = r in steps of 4): the A factors r x r row-major in global memory, the B
// factors r x nc with rows ldb apart. epi(t, c, p1, p2) receives each product entry with t < r, c < nc.
template <bool TWO, class Epi>
__device__ __forceinline__ void rr_gemm(const double *A1, const double *B1, const double *A2, const double *B2, int ldb, int r, int nc, int wv,
                                        int nwv, int lane, Epi epi) {
  const int tr = (r + 15) >> 4, tcn = (nc + 15) >> 4, nks = (r + 3) >> 2;
  for (int q = wv; q < tr * tcn; q += nwv) {
    const int t0 = 16 * (q / tcn), c0 = 16 * (q % tcn);
    const int ta = t0 + (lane & 15), cb = c0 + (lane & 15);
    v4d p1 = {0.0, 0.0, 0.0, 0.0}, p2 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int ks = 0; ks < nks; ++ks) {
      const int k = 4 * ks + (lane >> 4);
      const bool ak = ta < r && k < r, bk = cb < nc && k < r;
      p1 = mfma16x16x4(ak ? A1[ta * r + k] : 0.0, bk ? B1[(size_t)k * ldb + cb] : 0.0, p1);
      if (TWO) p2 = mfma16x16x4(ak ? A2[ta * r + k] : 0.0, bk ? B2[(size_t)k * ldb + cb] : 0.0, p2);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int t = t0 + 4 * v + (lane >> 4);
      if (t < r && cb < nc) epi(t, cb, p1[v], p2[v]);
    }
  }
}

// Dyson solve of one block per workgroup (512 threads), in the coefficient form of
// dyson_it_ppsc::solve (dlr_dyson_ppsc.hpp:135-168):
//   sigc = vals2coefs(Sigma_b), ggs = -(convolve_TO(beta, g0c, sigc) + eta G0_b), f = vals2coefs(ggs)
//   (I + convmat(f)) G = G0  with  convmat(f) = Mc cf2it^{-1}  is solved as
//   (cf2it x I + Mc) g = G0,  G = (cf2it x I) g,
// where Mc[(t,i),(k',i')] = beta ( -cf2it(t,k') (W f)_k'[i,i'] + tcf2it(t,k') f_k'[i,i']
//                                  - sum_k cf2it(t,k) f_k[i,i'] W(k,k') )
// maps DLR coefficients to values. The same equation without the inverse of cf2it: its rounding
// sensitivity is 2e-16 against 6e-16 for the transposed-solve form and 4e-9 for an explicit inverse.
// The N x N system (N = r n, padded to a multiple of 16 with an identity block) is factorised by a
// right-looking blocked LU with partial pivoting: 16-column panels factorised by one wave in LDS,
// then the trailing update A22 -= L21 U12 on FP64 MFMA tiles. The n right-hand sides (columns of
// G0) are solved in LDS by blocked substitution, G = (cf2it x I) g is formed and hermitised
// (block_sparse_solver.py:579-587). Sequential work (column solves with the DLR factors, panel
// pivoting, diagonal-block substitution) runs on LDS copies; the arithmetic is the same as on the
// global arrays.
__global__ void __launch_bounds__(kDysonThreads) dyson_kernel(DysonBatch Bt) {
  __shared__ double s_lu[kMaxRank * kMaxRank];  // LU factors of cf2it
  __shared__ int s_piv[kMaxRank];
  __shared__ double s_dg[16 * 16];               // diagonal block of the substitution
  __shared__ int s_ip[kDysonMaxNp];              // row interchanges of the factorisation
  // dynamic LDS: [the system (np_max^2 doubles), when np_max <= kLdsSysMax] then np_max x 16 doubles
  // for the LU panel ((Np - k0) rows x 16 columns), the r x n^2 columns of a vals2coefs, or the
  // Np x n right-hand sides
  extern __shared__ double lds_dyn[];
  const bool sys_lds = Bt.np_max <= kLdsSysMax;
  double *Pn = lds_dyn + (sys_lds ? (size_t)Bt.np_max * Bt.np_max : 0);
  const int pidx = find_seg(Bt.prefix, Bt.nprobs, blockIdx.x);
  const DysonArgs &A = Bt.probs[pidx];
  const int b = blockIdx.x - Bt.prefix[pidx], n = A.dims[b], r = A.r, nn = n * n, N = r * n;
  if (A.large && A.large[b]) return;  // the whole workgroup: this block's system is solved device-wide
  const int Np = (N + 15) & ~15;
  const double eta = *A.eta;
  const int tid = threadIdx.x, NT = blockDim.x, lane = tid & 63, wv = tid >> 6, NWV = NT >> 6;
  const size_t go = A.goff[b];
  const double *G0 = A.G0 + go, *g0c = A.g0c + go, *Wg0c = A.Wg0c + go;
  double *ggs = A.ggs + go, *wf = A.wf + go;
  double *S = sys_lds ? lds_dyn : A.sys + A.soff[b];  // Np x Np system, then its LU factors
  double *B = A.rhs + A.roff[b];
  int *ip = A.ipiv + A.poff[b];
  if (Bt.mode == 2) {  // the system was factorised across workgroups: its pivots
    for (int k = tid; k < Np; k += NT) s_ip[k] = ip[k];
    __syncthreads();
  } else {
  for (int x = tid; x < r * r; x += NT) s_lu[x] = A.lu[x];
  for (int x = tid; x < r; x += NT) s_piv[x] = A.piv[x];
  // ---- sigc = vals2coefs(Sigma_b), the columns solved in LDS (r n^2 <= r n 16 <= np_max 16)
  double *X = Pn;
  for (int x = tid; x < r * nn; x += NT) X[x] = A.sigma[go + x];
  __syncthreads();
  for (int c = wv; c < nn; c += NWV) {  // one wave per column
    const double v = lu_solve_wave(s_lu, s_piv, r, lane < r ? X[(size_t)lane * nn + c] : 0.0, lane);
    if (lane < r) X[(size_t)lane * nn + c] = v;
  }
  __syncthreads();
  // ---- W sigc -> wf (scratch), X holds sigc
  rr_gemm<false>(A.W, X, nullptr, nullptr, nn, r, nn, wv, NWV, lane,
                 [&](int t, int c, double p1, double) { wf[(size_t)t * nn + c] = p1; });
  __syncthreads();
  // ---- node products: c = -(Wg0c sigc + g0c Wsig) -> ggs (temporarily), d = g0c sigc -> B (scratch)
  for (int x = tid; x < r * nn; x += NT) {
    const int t = x / nn, i = (x % nn) / n, j = x % n;
    const double *f = g0c + (size_t)t * nn, *wfk = Wg0c + (size_t)t * nn, *g = X + (size_t)t * nn, *wg = wf + (size_t)t * nn;
    double s1 = 0.0, s2 = 0.0;
    for (int l = 0; l < n; ++l) {
      s1 += wfk[i * n + l] * g[l * n + j] + f[i * n + l] * wg[l * n + j];
      s2 += f[i * n + l] * g[l * n + j];
    }
    ggs[x] = -s1;
    B[x] = s2;
  }
  __syncthreads();
  // ---- ggs = -(beta (cf2it c + tcf2it d) + eta G0) -> wf, then copied to ggs
  rr_gemm<true>(A.cf2it, ggs, A.tcf2it, B, nn, r, nn, wv, NWV, lane, [&](int t, int c, double h1, double h2) {
    wf[(size_t)t * nn + c] = -(A.beta * (h1 + h2) + eta * G0[(size_t)t * nn + c]);
  });
  __syncthreads();
  for (int x = tid; x < r * nn; x += NT) ggs[x] = wf[x];
  __syncthreads();
  // ---- f = vals2coefs(ggs) (solved in LDS), W f -> wf
  for (int x = tid; x < r * nn; x += NT) X[x] = ggs[x];
  __syncthreads();
  for (int c = wv; c < nn; c += NWV) {
    const double v = lu_solve_wave(s_lu, s_piv, r, lane < r ? X[(size_t)lane * nn + c] : 0.0, lane);
    if (lane < r) X[(size_t)lane * nn + c] = v;
  }
  __syncthreads();
  rr_gemm<false>(A.W, X, nullptr, nullptr, nn, r, nn, wv, NWV, lane,
                 [&](int t, int c, double p1, double) { wf[(size_t)t * nn + c] = p1; });
  for (int x = tid; x < r * nn; x += NT) ggs[x] = X[x];
  // the LU factors are done with: cf2it takes their LDS copy for the system (X keeps f)
  double *s_cf = s_lu;
  __syncthreads();
  for (int x = tid; x < r * r; x += NT) s_cf[x] = A.cf2it[x];
  __syncthreads();
  // ---- system cf2it x I + Mc, padded with identity to Np x Np. The sum over k of Mc is, for each
  // orbital pair (i, i'), the product (cf2it diag(f_{i i'})) W: 16 x 16 MFMA tiles over (t, k').
  for (int x = tid; x < (Np - N) * Np; x += NT) {  // padding rows
    const int row = N + x / Np, col = x % Np;
    S[(size_t)row * Np + col] = (row == col) ? 1.0 : 0.0;
  }
  for (int x = tid; x < N * (Np - N); x += NT) S[(size_t)(x / (Np - N)) * Np + N + x % (Np - N)] = 0.0;  // padding columns
  {
    // a wave takes one (t, k') tile and a run of orbital pairs: the tile's W fragments and tcf2it
    // entries are read once for the run (runs sized so that every wave has a task)
    const int RBd = (r + 15) >> 4, nks = (r + 3) >> 2, ntile = RBd * RBd;
    const int run = max(1, ntile * nn / NWV), nrun = (nn + run - 1) / run;
    for (int q = wv; q < ntile * nrun; q += NWV) {
      const int tile = q % ntile, c0 = (q / ntile) * run, c1 = min(nn, c0 + run);
      const int t0 = 16 * (tile / RBd), kp0 = 16 * (tile % RBd);
      const int ta = t0 + (lane & 15), kpb = kp0 + (lane & 15);
      double bw[kMaxRank / 4], tc[4];
#pragma unroll
      for (int ks = 0; ks < kMaxRank / 4; ++ks) {
        const int k = 4 * ks + (lane >> 4);
        bw[ks] = (ks < nks && k < r && kpb < r) ? A.W[k * r + kpb] : 0.0;
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int t = t0 + 4 * v + (lane >> 4);
        tc[v] = (t < r && kpb < r) ? A.tcf2it[t * r + kpb] : 0.0;
      }
      for (int c = c0; c < c1; ++c) {
        const int i = c / n, ipp = c % n;
        v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < kMaxRank / 4; ++ks) {
          if (ks < nks) {
            const int k = 4 * ks + (lane >> 4);
            const double a = (ta < r && k < r) ? s_cf[ta * r + k] * X[(size_t)k * nn + c] : 0.0;
            acc = mfma16x16x4(a, bw[ks], acc);
          }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int t = t0 + 4 * v + (lane >> 4), kp = kpb;
          if (t < r && kp < r) {
            const double sv = -s_cf[t * r + kp] * wf[(size_t)kp * nn + c] + tc[v] * X[(size_t)kp * nn + c] - acc[v];
            S[(size_t)(t * n + i) * Np + kp * n + ipp] = A.beta * sv + (i == ipp ? s_cf[t * r + kp] : 0.0);
          }
        }
      }
    }
  }
  __syncthreads();
  if (Bt.mode == 1) return;  // the workgroup-wide LU launches follow (lu_panel_kernel, lu_update_kernel)
  // ---- blocked LU with partial pivoting. Each 16-column panel is factorised in LDS (rows k0..Np-1),
  // written back, its row interchanges applied to the other columns together with U12 = L11^{-1} A12
  // (one column per thread), then A22 -= L21 U12 on MFMA tiles with L21 read from LDS.
  for (int k0 = 0; k0 < Np; k0 += 16) {
    const int m = Np - k0;
    for (int x = tid; x < m * 16; x += NT) Pn[x] = S[(size_t)(k0 + (x >> 4)) * Np + k0 + (x & 15)];
    __syncthreads();
    if (wv == 0) {  // one wave: no workgroup barrier per column
      if (m <= 64 * kWavePanelSlots) panel_factor_regs(Pn, m, k0, Np, s_ip + k0, lane);
      else panel_factor(Pn, m, k0, Np, s_ip + k0, lane);
    }
    __syncthreads();
    for (int x = tid; x < m * 16; x += NT) S[(size_t)(k0 + (x >> 4)) * Np + k0 + (x & 15)] = Pn[x];
    // the panel's row interchanges on every other column (in order, per column), and U12
    for (int c = tid; c < Np - 16; c += NT) panel_column(S, Np, k0, c < k0 ? c : c + 16, s_ip + k0, Pn);
    __syncthreads();
    const int c1 = k0 + 16;
    if (c1 >= Np) break;
    // A22 -= L21 U12 on 16 x 16 MFMA tiles, one tile per wave at a time, L21 read from LDS
    const int nt = (Np - c1) >> 4;
    for (int q = wv; q < nt * nt; q += NWV) {
      const int ti = c1 + 16 * (q / nt), tj = c1 + 16 * (q % nt);
      trailing_tile(S, Np, k0, ti, tj, Pn + (size_t)(ti - k0) * 16, 16, lane);
    }
    __syncthreads();
  }
  }
  // ---- right-hand sides in LDS (Np n <= np_max 16): Y[(t,i)][j] = G0[t][i][j], padding rows zero;
  // row interchanges in order
  double *Y = Pn;
  for (int x = tid; x < Np * n; x += NT) Y[x] = (x < N * n) ? G0[x] : 0.0;
  __syncthreads();
  for (int col = tid; col < n; col += NT)
    for (int k = 0; k < Np; ++k)
      if (s_ip[k] != k) {
        const double a = Y[(size_t)k * n + col];
        Y[(size_t)k * n + col] = Y[(size_t)s_ip[k] * n + col];
        Y[(size_t)s_ip[k] * n + col] = a;
      }
  __syncthreads();
  // forward substitution with the unit lower factor, 16 rows at a time (diagonal block in LDS)
  for (int k0 = 0; k0 < Np; k0 += 16) {
    for (int x = tid; x < 256; x += NT) s_dg[x] = S[(size_t)(k0 + (x >> 4)) * Np + k0 + (x & 15)];
    __syncthreads();
    for (int col = wv; col < n; col += NWV) {  // one wave per column, lane q holding row k0 + q
      double v = lane < 16 ? Y[(size_t)(k0 + lane) * n + col] : 0.0;
#pragma unroll
      for (int q0 = 0; q0 < 15; ++q0) {
        const double l = s_dg[(lane & 15) * 16 + q0];
        const double u = v - l * lane_value(v, q0);
        v = (lane > q0 && lane < 16) ? u : v;
      }
      if (lane < 16) Y[(size_t)(k0 + lane) * n + col] = v;
    }
    __syncthreads();
    for (int i = k0 + 16 + tid; i < Np; i += NT) {  // one row per thread, its 16 factor entries read together
      double l[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) l[q] = S[(size_t)i * Np + k0 + q];
      for (int col = 0; col < n; ++col) {
        double s = Y[(size_t)i * n + col];
#pragma unroll
        for (int q = 0; q < 16; ++q) s -= l[q] * Y[(size_t)(k0 + q) * n + col];
        Y[(size_t)i * n + col] = s;
      }
    }
    __syncthreads();
  }
  // backward substitution with the upper factor
  for (int k0 = Np - 16; k0 >= 0; k0 -= 16) {
    for (int x = tid; x < 256; x += NT) s_dg[x] = S[(size_t)(k0 + (x >> 4)) * Np + k0 + (x & 15)];
    __syncthreads();
    for (int col = wv; col < n; col += NWV) {  // one wave per column, from the last row up
      double v = lane < 16 ? Y[(size_t)(k0 + lane) * n + col] : 0.0;
#pragma unroll
      for (int q = 15; q >= 0; --q) {
        const double dq = s_dg[q * 16 + q], l = s_dg[(lane & 15) * 16 + q];
        if (lane == q) v = v / dq;
        const double u = v - l * lane_value(v, q);
        v = lane < q ? u : v;
      }
      if (lane < 16) Y[(size_t)(k0 + lane) * n + col] = v;
    }
    __syncthreads();
    for (int i = tid; i < k0; i += NT) {
      double u[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) u[q] = S[(size_t)i * Np + k0 + q];
      for (int col = 0; col < n; ++col) {
        double s = Y[(size_t)i * n + col];
#pragma unroll
        for (int q = 0; q < 16; ++q) s -= u[q] * Y[(size_t)(k0 + q) * n + col];
        Y[(size_t)i * n + col] = s;
      }
    }
    __syncthreads();
  }
  for (int k = tid; k < Np; k += NT) ip[k] = s_ip[k];  // the pivots, kept for inspection
  // ---- G[t][i][j] = sum_k cf2it(t,k) g[(k,i)][j], hermitised
  double *Gout = A.gout + go;
  rr_gemm<false>(A.cf2it, Y, nullptr, nullptr, nn, r, nn, wv, NWV, lane,
                 [&](int t, int c, double p1, double) { Gout[(size_t)t * nn + c] = p1; });
  __syncthreads();
  for (int x = tid; x < r * nn; x += NT) {
    const int t = x / nn, i = (x / n) % n, j = x % n;
    if (i < j) {
      const double a = Gout[(size_t)t * nn + i * n + j], c = Gout[(size_t)t * nn + j * n + i];
      const double h = 0.5 * (a + c);
      Gout[(size_t)t * nn + i * n + j] = h;
      Gout[(size_t)t * nn + j * n + i] = h;
    }
  }
}

constexpr int kPanelRows = (kDysonMaxNp + 255) / 256;  // panel rows per thread of lu_panel_kernel

// Column JJ of a panel held in registers by lu_panel_kernel (the row loaded as row i sits in thread
// i mod 256, slot i / 256, and pos[k] is its current position): the pivot (row of maximal modulus at
// the smallest position: DPP reduction per wave, then over the four waves in a fixed order), the
// interchange (two positions swapped) and the update of every row below, in the arithmetic of
// panel_factor. Each wave publishes its candidate row with its candidate before the one workgroup
// barrier of the column; the LDS slots alternate between columns, so a column's writes never meet
// the previous column's reads. Column index and slots are compile-time constants, so the rows stay
// in registers.
template <int JJ>
__device__ __forceinline__ void panel_step(double (&v)[kPanelRows][16], int (&pos)[kPanelRows], int m, int k0, int Np, int tid,
                                           double (*s_cv)[4], int (*s_ci)[4], double (*s_crow)[4][16], int *ipk) {
  constexpr int bs = JJ & 1;
  const int lane = tid & 63, wv = tid >> 6;
  double best = -1.0;
  int bp = Np;
#pragma unroll
  for (int k = 0; k < kPanelRows; ++k)
    if (pos[k] >= JJ && pos[k] < m) {
      const double a = fabs(v[k][JJ]);
      if (a > best || (a == best && pos[k] < bp)) { best = a; bp = pos[k]; }
    }
  wave_argmax(best, bp, Np);
  if (lane == 0) {
    s_cv[bs][wv] = best;
    s_ci[bs][wv] = bp;
  }
#pragma unroll
  for (int k = 0; k < kPanelRows; ++k)
    if (pos[k] == bp) {
#pragma unroll
      for (int c = 0; c < 16; ++c) s_crow[bs][wv][c] = v[k][c];
    }
  __syncthreads();
  double bv = -1.0;
  int p = Np, pw = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const double cv = s_cv[bs][w];
    const int ci = s_ci[bs][w];
    if (cv > bv || (cv == bv && ci < p)) { bv = cv; p = ci; pw = w; }
  }
  if (tid == 0) ipk[JJ] = k0 + p;
  double prow[16];
#pragma unroll
  for (int c = JJ; c < 16; ++c) prow[c] = s_crow[bs][pw][c];
#pragma unroll
  for (int k = 0; k < kPanelRows; ++k) pos[k] = pos[k] == p ? JJ : (pos[k] == JJ ? p : pos[k]);
  const double pv = prow[JJ];
#pragma unroll
  for (int k = 0; k < kPanelRows; ++k)
    if (pos[k] > JJ && pos[k] < m) {
      const double l = v[k][JJ] / pv;
      v[k][JJ] = l;
#pragma unroll
      for (int c = JJ + 1; c < 16; ++c) v[k][c] -= l * prow[c];
    }
}

template <int JJ>
__device__ __forceinline__ void panel_steps(double (&v)[kPanelRows][16], int (&pos)[kPanelRows], int m, int k0, int Np, int tid,
                                            double (*s_cv)[4], int (*s_ci)[4], double (*s_crow)[4][16], int *ipk) {
  if constexpr (JJ < 16) {
    panel_step<JJ>(v, pos, m, k0, Np, tid, s_cv, s_ci, s_crow, ipk);
    panel_steps<JJ + 1>(v, pos, m, k0, Np, tid, s_cv, s_ci, s_crow, ipk);
  }
}

// Panel k0 of every block's system (assembled in global memory by dyson_kernel in mode 1), one
// workgroup of four waves per block with the panel's rows in registers, one workgroup barrier per
// column (panel_step). Blocks whose system is smaller than k0 + 16 rows, or solved by the
// large-block path, exit.
__global__ void __launch_bounds__(256) lu_panel_kernel(DysonBatch Bt, int k0) {
  __shared__ double s_cv[2][4], s_crow[2][4][16];
  __shared__ int s_ci[2][4], ipk[16];
  const int pidx = find_seg(Bt.prefix, Bt.nprobs, blockIdx.x);
  const DysonArgs &A = Bt.probs[pidx];
  const int b = blockIdx.x - Bt.prefix[pidx], Np = (A.r * A.dims[b] + 15) & ~15;
  if ((A.large && A.large[b]) || k0 >= Np) return;
  double *S = A.sys + A.soff[b];
  const int tid = threadIdx.x, m = Np - k0;
  double v[kPanelRows][16];
  int pos[kPanelRows];
#pragma unroll
  for (int k = 0; k < kPanelRows; ++k) {
    const int i = tid + 256 * k;
    pos[k] = i;
#pragma unroll
    for (int c = 0; c < 16; ++c) v[k][c] = i < m ? S[(size_t)(k0 + i) * Np + k0 + c] : 0.0;
  }
  panel_steps<0>(v, pos, m, k0, Np, tid, s_cv, s_ci, s_crow, ipk);
  __syncthreads();  // every thread has read its rows before any is written to its final position
#pragma unroll
  for (int k = 0; k < kPanelRows; ++k)
    if (pos[k] < m) {
#pragma unroll
      for (int c = 0; c < 16; ++c) S[(size_t)(k0 + pos[k]) * Np + k0 + c] = v[k][c];
    }
  __syncthreads();
  if (tid < 16) A.ipiv[A.poff[b] + k0 + tid] = ipk[tid];
}

// After panel k0: one workgroup per 16-column strip (blockIdx.x) of block blockIdx.y, every strip
// but the panel's. The strip's columns take the panel's row interchanges, applied as one gather of
// the (at most 32) rows they touch, and, right of the panel, U12 = L11^{-1} A12 (one column per
// thread); then its column of trailing tiles A22 -= L21 U12 (one tile per wave at a time, L21 read
// from the factorised panel in global memory). The same arithmetic in the same order as the
// single-workgroup LU of dyson_kernel (panel_column, trailing_tile).
__global__ void __launch_bounds__(256) lu_update_kernel(DysonBatch Bt, int k0) {
  __shared__ double L11[256], U[256];
  __shared__ int ipk[16], rows[32], src[32], nrows;
  const int pidx = find_seg(Bt.prefix, Bt.nprobs, blockIdx.y);
  const DysonArgs &A = Bt.probs[pidx];
  const int b = blockIdx.y - Bt.prefix[pidx], Np = (A.r * A.dims[b] + 15) & ~15;
  if ((A.large && A.large[b]) || k0 >= Np) return;
  const int j = blockIdx.x, nstrips = Np / 16 - 1;
  if (j >= nstrips) return;
  const int col0 = 16 * j < k0 ? 16 * j : 16 * (j + 1);
  double *S = A.sys + A.soff[b];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  L11[tid] = S[(size_t)(k0 + (tid >> 4)) * Np + k0 + (tid & 15)];
  if (tid < 16) ipk[tid] = A.ipiv[A.poff[b] + k0 + tid];
  __syncthreads();
  if (wv == 0) {
    // the panel's interchanges, in order, as a permutation of the rows they touch, one row per lane
    // (lanes 0..15 rows k0..k0+15, then the other pivot rows): after them row rows[x] holds what row
    // src[x] held before
    int row = lane < 16 ? k0 + lane : -1, cnt = 16;
    for (int q = 0; q < 16; ++q) {
      const int pr = ipk[q];
      if (__ballot(row == pr) == 0) {
        if (lane == cnt) row = pr;
        ++cnt;
      }
    }
    int sr = row;
    for (int q = 0; q < 16; ++q) {
      const int pr = ipk[q];
      if (pr == k0 + q) continue;
      const int xb = __ffsll((unsigned long long)__ballot(row == pr)) - 1;
      const int sa = __shfl(sr, q, 64), sb = __shfl(sr, xb, 64);
      if (lane == q) sr = sb;
      if (lane == xb) sr = sa;
    }
    if (lane < 32) {
      rows[lane] = row;
      src[lane] = sr;
    }
    if (lane == 0) nrows = cnt;
  }
  __syncthreads();
  {  // the strip's columns gathered through the permutation (all reads before any write)
    const int c = tid & 15, x0 = tid >> 4, cnt = nrows;
    double v0 = 0.0, v1 = 0.0;
    if (x0 < cnt) v0 = S[(size_t)src[x0] * Np + col0 + c];
    if (x0 + 16 < cnt) v1 = S[(size_t)src[x0 + 16] * Np + col0 + c];
    U[x0 * 16 + c] = v0;  // rows k0..k0+15 of the strip after the interchanges
    __syncthreads();
    if (x0 < cnt && src[x0] != rows[x0]) S[(size_t)rows[x0] * Np + col0 + c] = v0;
    if (x0 + 16 < cnt && src[x0 + 16] != rows[x0 + 16]) S[(size_t)rows[x0 + 16] * Np + col0 + c] = v1;
  }
  __syncthreads();
  if (col0 >= k0 + 16 && tid < 16) {  // U12 = L11^{-1} A12, one column per thread, as in panel_column
    double x[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) x[q] = U[q * 16 + tid];
#pragma unroll
    for (int q = 1; q < 16; ++q) {
      double sacc = x[q];
#pragma unroll
      for (int q0 = 0; q0 < 16; ++q0)
        if (q0 < q) sacc -= L11[q * 16 + q0] * x[q0];
      x[q] = sacc;
    }
#pragma unroll
    for (int q = 1; q < 16; ++q) S[(size_t)(k0 + q) * Np + col0 + tid] = x[q];
  }
  __syncthreads();
  const int c1 = k0 + 16;
  if (col0 < c1) return;
  for (int ti = c1 + 16 * wv; ti < Np; ti += 64) trailing_tile(S, Np, k0, ti, col0, S + (size_t)ti * Np + k0, Np, lane);
}

// Per problem (one workgroup of four waves each): Z = -Tr G_new(beta) = -sum_t zw[t] Tr G_new(t).
// Wave w sums the diagonal entries of the blocks b = w, w + 4, ... per node (lane t), the four wave
// sums are added per node in wave order and the node terms in node order by one lane, so the sum has
// a fixed order; deta = ln|Z| / beta, eta += deta; the difference accumulator is cleared for the
// update kernel that follows.
__global__ void __launch_bounds__(256) trace_kernel(IterBatch Bt) {
  __shared__ double part[4][64];
  const IterArgs &A = Bt.probs[blockIdx.x];
  const int t = threadIdx.x & 63, w = threadIdx.x >> 6, r = A.r;
  double s = 0.0;
  if (t < r)
    for (int b = w; b < A.nb; b += 4) {
      const int n = A.dims[b];
      const double *g = A.Gnew + A.goff[b] + (size_t)t * n * n;
#pragma unroll 4
      for (int i = 0; i < n; ++i) s += g[i * n + i];
    }
  part[w][t] = s;
  __syncthreads();
  if (threadIdx.x < 64) part[0][t] = t < r ? A.zw[t] * (((part[0][t] + part[1][t]) + part[2][t]) + part[3][t]) : 0.0;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tr = 0.0;
    for (int q = 0; q < r; ++q) tr += part[0][q];
    const double Z = -tr;
    const double deta = log(fabs(Z)) / A.beta;
    A.status[kStZ] = Z;
    A.status[kStDeta] = deta;
    A.status[kStEta] += deta;
    A.status[kStDiff] = 0.0;
  }
}

// G_new *= exp(-tau deta); status[diff] = max |G_new - G| (non-negative doubles order like their bits)
__global__ void __launch_bounds__(256) scale_diff_kernel(IterBatch Bt) {
  const IterArgs &A = Bt.probs[blockIdx.y];
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (x < A.total) {
    int b = 0;
    while (b + 1 < A.nb && A.goff[b + 1] <= x) ++b;
    const int n = A.dims[b];
    const int t = (int)((x - A.goff[b]) / ((int64_t)n * n));
    const double v = A.Gnew[x] * exp(-A.tau[t] * A.status[kStDeta]);
    A.Gnew[x] = v;
    d = fabs(v - A.G[x]);
  }
  if (blockIdx.x * blockDim.x < A.total)
    atomicMax(reinterpret_cast<unsigned long long *>(A.status + kStDiff), (unsigned long long)__double_as_longlong(d));
}

// Per (problem, block): converged (diff < tol): G stays (the solver stops without mixing), otherwise
// G = mix G_new + (1 - mix) G; then max |G| of the block, in a fixed order. The first block's
// workgroup records the convergence flag.
__global__ void __launch_bounds__(256) mix_max_kernel(IterBatch Bt) {
  __shared__ double part[256];
  const int pidx = find_seg(Bt.prefix, Bt.nprobs, blockIdx.x);
  const IterArgs &A = Bt.probs[pidx];
  const int b = blockIdx.x - Bt.prefix[pidx];
  const bool conv = A.status[kStDiff] < A.tol;
  double m = 0.0;
  for (int64_t x = A.goff[b] + threadIdx.x; x < A.gend[b]; x += blockDim.x) {
    double g = A.G[x];
    if (!conv) {
      g = A.mix * A.Gnew[x] + (1.0 - A.mix) * A.G[x];
      A.G[x] = g;
    }
    m = fmax(m, fabs(g));
  }
  part[threadIdx.x] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < (int)blockDim.x; ++q) t = fmax(t, part[q]);
    A.status[kStBmax + b] = t;
    if (b == 0) A.status[kStConv] = conv ? 1.0 : 0.0;
  }
}

__global__ void fill_kernel(int64_t n, double v, double *p) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) p[x] = v;
}

// host-callable launchers
hipError_t launch_dyson(const DysonBatch &B, hipStream_t s) {
  static bool ready = false;  // attributes set once, outside any stream capture after the first launch
  if (!ready) {
    const size_t most = std::max<size_t>((size_t)kDysonMaxNp * 16, (size_t)kLdsSysMax * (kLdsSysMax + 16));
    hipError_t e = hipFuncSetAttribute((const void *)dyson_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)(sizeof(double) * most));
    if (e != hipSuccess) return e;
    ready = true;
  }
  if (B.total <= 0) return hipSuccess;
  if (B.np_max > kDysonMaxNp) return hipErrorInvalidValue;
  const size_t np = (size_t)B.np_max;
  const size_t sh = sizeof(double) * ((np <= (size_t)kLdsSysMax ? np * np : 0) + np * 16);
  if (B.np_max <= kDysonMultiNp) {
    DysonBatch one = B;
    one.mode = 0;
    hipLaunchKernelGGL(dyson_kernel, dim3(B.total), dim3(kDysonThreads), sh, s, one);
    return hipGetLastError();
  }
  // systems in global memory, factorised across workgroups panel by panel
  DysonBatch pre = B, post = B;
  pre.mode = 1;
  post.mode = 2;
  hipLaunchKernelGGL(dyson_kernel, dim3(B.total), dim3(kDysonThreads), sh, s, pre);
  for (int k0 = 0; k0 < B.np_max; k0 += 16) {
    hipLaunchKernelGGL(lu_panel_kernel, dim3(B.total), dim3(256), 0, s, B, k0);
    hipLaunchKernelGGL(lu_update_kernel, dim3(B.np_max / 16 - 1, B.total), dim3(256), 0, s, B, k0);
  }
  hipLaunchKernelGGL(dyson_kernel, dim3(B.total), dim3(kDysonThreads), sh, s, post);
  return hipGetLastError();
}

hipError_t launch_iterate(const IterBatch &B, hipStream_t s) {
  if (B.nprobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(trace_kernel, dim3(B.nprobs), dim3(256), 0, s, B);
  const dim3 grid((unsigned)((B.max_total + 255) / 256), B.nprobs);
  hipLaunchKernelGGL(scale_diff_kernel, grid, dim3(256), 0, s, B);
  return hipGetLastError();
}

hipError_t launch_block_max(const IterBatch &B, hipStream_t s) {
  if (B.total_blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(mix_max_kernel, dim3(B.total_blocks), dim3(256), 0, s, B);
  return hipGetLastError();
}

hipError_t launch_fill(int64_t n, double v, double *p, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, v, p);
  return hipGetLastError();
}

}  // namespace xca::dev

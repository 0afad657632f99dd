// HIP kernels of the pseudo-particle self-energy engine (gfx950, FP64 throughout).
//
// Layout of one tile in LDS: two slabs of RP x LDN doubles (RP = 16 * RB padded DLR rank; rows are
// DLR nodes, columns are (item, row, column) triples of the per-item state T(t) for up to LDN
// columns). Items of one tile share a program (one block path and rule class) and differ by their
// pole tuple. Every node-space transform is an FP64 MFMA GEMM (v_mfma_f64_16x16x4f64) with the
// DLR operator as the A operand (pre-packed in lane order in global memory) and the tile as the B
// operand (LDS); each wave owns one 16-column block of the tile.
//
// Edge integration h = int_0^t GK(t-s) T(s) ds follows the time-ordered DLR convolution:
//   g  = vals2coefs(T)   : (L^{-1} P) T by MFMA, then back substitution with U in 4-row steps
//                          (4x4 diagonal-block inverses, rank-4 MFMA trailing updates) —
//                          backward stable, unlike a product with the explicit inverse of cf2it;
//   Wg = W g             : MFMA;
//   c  = -(Wf o g + f o Wg),  d = f o g   (per node, per item: small products, vector ALU);
//   h  = beta (cf2it c + tcf2it d)        : one MFMA GEMM with K = 2 RP.
// f = vals2coefs(GK) and W f come from edge tables computed by tab_kernel with the same routines.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "engine.hpp"

namespace xca::dev {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kMaxB = 12;   // largest block handled by the tile kernels
#ifndef XCA_GEMM_UNROLL
#define XCA_GEMM_UNROLL 2
#endif

__device__ __forceinline__ v4d mfma_f64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// acc[rb] += sum_{ks < NK} A[rb][koff + ks] * S[4 ks + (lane >> 4)][c0 + (lane & 15)]
// A fragments packed as Afrag[(rb * ldk + ks) * 64 + lane]; the next step's A and B fragments are
// loaded before the current step's MFMAs. NK = ceil(r / 4): the zero padding of the DLR rank up to
// the 16-row blocks is skipped in the contraction.
template <int RB>
__device__ __forceinline__ void gemm_acc(const double *__restrict__ Afrag, int ldk, int koff, int NK, const double *S, int ldn, int c0,
                                         v4d (&acc)[RB], int lane) {
  const int krow = lane >> 4, col = c0 + (lane & 15);
  const double *Ap = Afrag + (size_t)koff * 64 + lane;
  double ac[RB], an[RB], bc, bn = 0.0;
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) { ac[rb] = Ap[(size_t)rb * ldk * 64]; an[rb] = 0.0; }
  bc = S[(size_t)krow * ldn + col];
#pragma unroll XCA_GEMM_UNROLL
  for (int ks = 0; ks < NK; ++ks) {
    if (ks + 1 < NK) {
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) an[rb] = Ap[((size_t)rb * ldk + ks + 1) * 64];
      bn = S[(size_t)(4 * (ks + 1) + krow) * ldn + col];
    }
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) acc[rb] = mfma_f64(ac[rb], bc, acc[rb]);
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) ac[rb] = an[rb];
    bc = bn;
  }
}

template <int RB>
__device__ __forceinline__ void zero_acc(v4d (&acc)[RB]) {
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) acc[rb] = v4d{0.0, 0.0, 0.0, 0.0};
}

// Accumulator layout of v_mfma_f64_16x16x4f64 (measured on gfx950): element (i, j) of the 16x16
// block sits in lane j + 16 (i % 4), register i / 4. Row of register v in this lane:
__device__ __forceinline__ int acc_row(int rb, int v, int lane) { return 16 * rb + 4 * v + (lane >> 4); }

// store accumulators to LDS, column c0 + (lane & 15). Rows at or beyond the DLR rank r are stored
// as exact zeros: the padding rows must stay zero for every later GEMM that contracts over them, and
// columns a tile does not use hold stale values that may have overflowed (0 * inf would be NaN).
template <int RB>
__device__ __forceinline__ void store_acc(const v4d (&acc)[RB], double *S, int ldn, int c0, int lane, int r) {
  const int col = c0 + (lane & 15);
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = acc_row(rb, v, lane);
      S[(size_t)row * ldn + col] = (row < r) ? acc[rb][v] : 0.0;
    }
}

// Row-partitioned GEMM: the wave computes row block rb of the product for NCB column blocks
// cb0 .. cb0 + NCB - 1: acc[c] += sum_{ks < NK} A[rb][koff + ks] * S[4 ks + (lane >> 4)][16 (cb0 + c) + (lane & 15)].
// The A fragments of the row block (global) are read in groups of four, the next group while the
// current one is used (four to seven steps ahead), so that their latency is paid about once per
// product however few column blocks are in use; B fragments come from LDS.
template <int NCB>
__device__ __forceinline__ void gemm_rows(const double *__restrict__ Afrag, int ldk, int rb, int koff, int NK, const double *S, int ldn,
                                          int cb0, int ncnt, v4d (&acc)[NCB], int lane) {
  if (ncnt <= 0) return;
  const int krow = lane >> 4, col = 16 * cb0 + (lane & 15);
  const double *Ap = Afrag + ((size_t)rb * ldk + koff) * 64 + lane;
  double ac[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) ac[j] = j < NK ? Ap[(size_t)j * 64] : 0.0;
  for (int k0 = 0; k0 < NK; k0 += 4) {
    double an[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) an[j] = k0 + 4 + j < NK ? Ap[(size_t)(k0 + 4 + j) * 64] : 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int ks = k0 + j;
      if (ks < NK) {
        double bf[NCB];
#pragma unroll
        for (int c = 0; c < NCB; ++c) bf[c] = (c < ncnt) ? S[(size_t)(4 * ks + krow) * ldn + col + 16 * c] : 0.0;
#pragma unroll
        for (int c = 0; c < NCB; ++c)
          if (c < ncnt) acc[c] = mfma_f64(ac[j], bf[c], acc[c]);
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) ac[j] = an[j];
  }
}

template <int NCB>
__device__ __forceinline__ void store_rows(const v4d (&acc)[NCB], double *S, int ldn, int rb, int cb0, int ncnt, int lane, int r) {
  const int col = 16 * cb0 + (lane & 15);
#pragma unroll
  for (int c = 0; c < NCB; ++c)
    if (c < ncnt)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = acc_row(rb, v, lane);
        S[(size_t)row * ldn + col + 16 * c] = (row < r) ? acc[c][v] : 0.0;
      }
}

template <int RB>
__device__ __forceinline__ void load_acc(v4d (&acc)[RB], const double *S, int ldn, int c0, int lane) {
  const int col = c0 + (lane & 15);
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[rb][v] = S[(size_t)acc_row(rb, v, lane) * ldn + col];
}

// Back substitution with U held in LDS: Ul = -U strictly above the diagonal, packed by rows
// (row i holds columns i+1 .. RP-1 at offset i (2 RP - i - 1) / 2), Dl = inverses of the 4x4
// diagonal blocks. Rows 4s..4s+3 of step s live in register s % 4 of row block s / 4, row 4s + k in
// lane group k, which is the B-operand layout of the rank-4 trailing update.
template <int RB>
__device__ __forceinline__ void usolve_lds(v4d (&acc)[RB], const double *Ul, const double *Dl, int r, int lane) {
  constexpr int RP = 16 * RB;
  const int k = lane >> 4, j = lane & 15;
  const int smax = (r - 1) >> 2;
#pragma unroll
  for (int s = 4 * RB - 1; s >= 0; --s) {
    if (s > smax) continue;
    const int rbI = s >> 2, v = s & 3;
    const double ys = acc[rbI][v];
    const double y0 = __shfl(ys, j, 64), y1 = __shfl(ys, j + 16, 64), y2 = __shfl(ys, j + 32, 64), y3 = __shfl(ys, j + 48, 64);
    const double *Di = Dl + 16 * s + 4 * k;
    const double x = Di[0] * y0 + Di[1] * y1 + Di[2] * y2 + Di[3] * y3;
    acc[rbI][v] = x;
    if (s > 0) {
      const int col = 4 * s + k;
#pragma unroll
      for (int rb = 0; rb <= rbI; ++rb) {
        const int row = 16 * rb + j;
        const double a = (row < 4 * s) ? Ul[row * (2 * RP - row - 1) / 2 + (col - row - 1)] : 0.0;
        acc[rb] = mfma_f64(a, x, acc[rb]);
      }
    }
  }
}

// Copy the packed U and the diagonal-block inverses into LDS.
template <int RB>
__device__ __forceinline__ void load_u(const DlrFrag &D, double *Ul, double *Dl, int tid, int nt) {
  constexpr int RP = 16 * RB, NU = RP * (RP - 1) / 2, ND = RP * 4;
  for (int x = tid; x < NU; x += nt) Ul[x] = D.Upk[x];
  for (int x = tid; x < ND; x += nt) Dl[x] = D.Dinv[x];
}

// Back substitution U x = y on the accumulators (rows of y in acc, one 16-column block per wave),
// 4 rows per step from the bottom. Rows 4s..4s+3 of step s live in register s % 4 of row block s / 4,
// row 4s + k in lane group k: exactly the B-operand layout of the rank-4 trailing update, so the
// solved rows feed the MFMA directly. The 4x4 diagonal block is applied through its inverse.
template <int RB>
__device__ __forceinline__ void usolve(v4d (&acc)[RB], const DlrFrag &D, int lane) {
  const int k = lane >> 4, j = lane & 15;
  const int smax = (D.r - 1) >> 2;  // steps above smax only see padding rows (identity, zero data)
#pragma unroll
  for (int s = 4 * RB - 1; s >= 0; --s) {
    if (s > smax) continue;
    const int rbI = s >> 2, v = s & 3;
    const double ys = acc[rbI][v];
    const double y0 = __shfl(ys, j, 64), y1 = __shfl(ys, j + 16, 64), y2 = __shfl(ys, j + 32, 64), y3 = __shfl(ys, j + 48, 64);
    const double *Di = D.Dinv + 16 * s + 4 * k;
    const double x = Di[0] * y0 + Di[1] * y1 + Di[2] * y2 + Di[3] * y3;
    acc[rbI][v] = x;
    if (s > 0) {
#pragma unroll
      for (int rb = 0; rb <= rbI; ++rb) acc[rb] = mfma_f64(D.AU[((size_t)s * RB + rb) * 64 + lane], x, acc[rb]);
    }
  }
}

// Coefficients of the RP x 16 column block c0 of slab S: acc <- U^{-1} L^{-1} P S
template <int RB>
__device__ __forceinline__ void vals2coefs_block(const DlrFrag &D, const double *S, int ldn, int c0, v4d (&acc)[RB], int lane) {
  zero_acc(acc);
  gemm_acc<RB>(D.LiP, 4 * RB, 0, (D.r + 3) >> 2, S, ldn, c0, acc, lane);
  usolve<RB>(acc, D, lane);
}

// -------------------------------------------------------------------------------------------
// Edge-function tables: f = vals2coefs(G_b K...) and W f for every entry of a table job.

// segment s of a batch holding global index k: the last s with prefix[s] <= k
__device__ __forceinline__ int find_segment(const int *prefix, int nsegs, int k) {
  int lo = 0, hi = nsegs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= k) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int RB, int NW>
__global__ void __launch_bounds__(64 * NW) tab_kernel(TabBatch B) {
  constexpr int RP = 16 * RB, NT = 64 * NW, LDN = 16 * NW + 8;
  extern __shared__ double lds[];
  double *SA = lds, *SB = lds + RP * LDN;
  double *Ul = SB + RP * LDN, *Dl = Ul + RP * (RP - 1) / 2;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int sidx = find_segment(B.prefix, B.nsegs, blockIdx.x);
  const TabSeg &T = B.segs[sidx];
  const DlrFrag D = T.D;
  const double *__restrict__ G = T.G, *__restrict__ kp = T.kp, *__restrict__ km = T.km;
  double *__restrict__ ftab = T.ftab;
  const int p = T.p;
  const DTabJob J = T.jobs[blockIdx.x - B.prefix[sidx]];
  const int r = D.r, b = J.b, bb = b * b, ncol = J.ne * bb;
  for (int x = tid; x < RP * LDN; x += NT) { SA[x] = 0.0; SB[x] = 0.0; }
  load_u<RB>(D, Ul, Dl, tid, NT);
  __syncthreads();
  // GK(t) for each entry: G_b (or its transpose) times the K factor of each line, four elements per
  // thread with all their reads in flight together
  const int n = r * ncol;
  for (int x0 = 0; x0 < n; x0 += 4 * NT) {
    double v[4], kf[4][kMaxLines];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int x = x0 + u * NT + tid;
      const bool act = x < n;
      const int t = act ? x / ncol : 0, c = act ? x % ncol : 0, e = c / bb, ij = c % bb;
      v[u] = act ? G[J.g_off + (size_t)t * bb + (J.tr ? (ij % b) * b + ij / b : ij)] : 0.0;
      int ee = J.e0 + e;
#pragma unroll
      for (int q = 0; q < kMaxLines; ++q) {
        const int l = ee % p;
        ee /= p;
        kf[u][q] = (act && q < J.nlines) ? (J.lsign[q] > 0 ? kp : km)[(size_t)l * r + t] : 1.0;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int x = x0 + u * NT + tid;
      if (x < n) {
        double y = v[u];
#pragma unroll
        for (int q = 0; q < kMaxLines; ++q)
          if (q < J.nlines) y *= kf[u][q];
        SA[(size_t)(x / ncol) * LDN + x % ncol] = y;
      }
    }
  }
  __syncthreads();
  // Row-partitioned GEMMs over the column blocks in use, as in the chain kernel: wave wv takes row
  // block wv % RB of the column blocks of its group wv / RB. Then g = U^{-1} y one column block per
  // wave (f written from the solved columns), and W g by the row-partitioned GEMM again.
  constexpr int NGRP = NW / RB, NCB = RB;
  static_assert(NW % RB == 0, "waves must tile the row blocks");
  const int rbw = wv % RB, grp = wv / RB;
  const int ncb = (ncol + 15) >> 4, cpg = (ncb + NGRP - 1) / NGRP, cbw = grp * cpg, ncw = min(cpg, ncb - cbw);
  const int nk = (r + 3) >> 2;
  {  // y = (L^{-1} P) GK -> SB
    v4d acc[NCB];
    zero_acc(acc);
    gemm_rows<NCB>(D.LiP, 4 * RB, rbw, 0, nk, SA, LDN, cbw, ncw, acc, lane);
    store_rows<NCB>(acc, SB, LDN, rbw, cbw, ncw, lane, r);
  }
  __syncthreads();
  for (int cb = wv; cb < ncb; cb += NW) {  // f = U^{-1} y in place, and to global
    v4d acc[RB];
    load_acc<RB>(acc, SB, LDN, 16 * cb, lane);
    usolve_lds<RB>(acc, Ul, Dl, r, lane);
    store_acc<RB>(acc, SB, LDN, 16 * cb, lane, r);
    const int col = 16 * cb + (lane & 15);
    if (col < ncol) {
      const int e = col / bb, ij = col % bb;
      double *out = ftab + J.out_off + (size_t)(J.e0 + e) * 2 * r * bb;
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int t = acc_row(rb, v, lane);
          if (t < r) out[(size_t)t * bb + ij] = acc[rb][v];
        }
    }
  }
  __syncthreads();
  {  // W f -> global
    v4d acc[NCB];
    zero_acc(acc);
    gemm_rows<NCB>(D.W, 4 * RB, rbw, 0, nk, SB, LDN, cbw, ncw, acc, lane);
#pragma unroll
    for (int c = 0; c < NCB; ++c) {
      const int col = 16 * (cbw + c) + (lane & 15);
      if (c < ncw && col < ncol) {
        const int e = col / bb, ij = col % bb;
        double *out = ftab + J.out_off + (size_t)(J.e0 + e) * 2 * r * bb + (size_t)r * bb;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int t = acc_row(rbw, v, lane);
          if (t < r) out[(size_t)t * bb + ij] = acc[c][v];
        }
      }
    }
  }
}

// -------------------------------------------------------------------------------------------
// Reflection t -> 1 - t of stored correlator outputs, in place: X <- E U^{-1} (L^{-1} P) X along
// the time index for every column of a job (values at the reflected nodes from the coefficients).
// Column c of a job is the r consecutive doubles at base + (c0 + c) r. One wave owns 16 columns;
// columns of different jobs never overlap.
template <int RB, int NW>
__global__ void __launch_bounds__(64 * NW) refl_kernel(ReflArgs A) {
  constexpr int RP = 16 * RB, NT = 64 * NW, LDN = 16 * NW + 8;
  extern __shared__ double lds[];
  double *SA = lds, *SB = lds + RP * LDN;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const DlrFrag D = A.D;
  const RJob J = A.jobs[blockIdx.x];
  const int r = D.r, ncol = J.ncol;
  for (int x = tid; x < RP * LDN; x += NT) { SA[x] = 0.0; SB[x] = 0.0; }
  __syncthreads();
  for (int x = tid; x < r * ncol; x += NT) {
    const int c = x / r, t = x % r;
    SA[(size_t)t * LDN + c] = A.store[J.base + (size_t)(J.c0 + c) * r + t];
  }
  __syncthreads();
  v4d acc[RB];
  const int c0 = 16 * wv;
  vals2coefs_block<RB>(D, SA, LDN, c0, acc, lane);
  store_acc<RB>(acc, SB, LDN, c0, lane, r);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  zero_acc(acc);
  gemm_acc<RB>(D.E, 4 * RB, 0, (r + 3) >> 2, SB, LDN, c0, acc, lane);
  const int c = c0 + (lane & 15);
  if (c < ncol) {
    double *out = A.store + J.base + (size_t)(J.c0 + c) * r;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int t = acc_row(rb, v, lane);
        if (t < r) out[t] = acc[rb][v];
      }
  }
}

// Large-block engine, node-space part of an edge (or of an edge table): for columns of X (item u:
// r rows of ld doubles at X + u xsi), G = vals2coefs(X) = U^{-1} (L^{-1} P) X and WG = W G, with the
// same fragments and substitution as tab_kernel. 16 NW columns per workgroup.
template <int RB, int NW>
__global__ void __launch_bounds__(64 * NW) wv2c_kernel(WV2C A) {
  constexpr int RP = 16 * RB, NT = 64 * NW, LDN = 16 * NW + 8;
  extern __shared__ double lds[];
  double *SA = lds, *SB = lds + RP * LDN;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const DlrFrag D = A.D;
  const int r = D.r;
  const int cblocks = (A.ncol + 16 * NW - 1) / (16 * NW);
  const int u = blockIdx.x / cblocks, cb0 = (blockIdx.x % cblocks) * 16 * NW;
  const int nc = min(16 * NW, A.ncol - cb0);
  for (int x = tid; x < RP * LDN; x += NT) { SA[x] = 0.0; SB[x] = 0.0; }
  __syncthreads();
  const double *X = A.X + (int64_t)u * A.xsi + cb0;
  for (int x = tid; x < r * nc; x += NT) {
    const int t = x / nc, c = x % nc;
    SA[(size_t)t * LDN + c] = X[(int64_t)t * A.ld + c];
  }
  __syncthreads();
  v4d acc[RB];
  const int c0 = 16 * wv;
  vals2coefs_block<RB>(D, SA, LDN, c0, acc, lane);
  store_acc<RB>(acc, SB, LDN, c0, lane, r);
  const int col = c0 + (lane & 15);
  if (col < nc) {
    double *g = A.G + (int64_t)u * A.gsi + cb0 + col;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int t = acc_row(rb, v, lane);
        if (t < r) g[(int64_t)t * A.ld] = acc[rb][v];
      }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  zero_acc(acc);
  gemm_acc<RB>(D.W, 4 * RB, 0, (r + 3) >> 2, SB, LDN, c0, acc, lane);
  if (col < nc) {
    double *wg = A.WG + (int64_t)u * A.wsi + cb0 + col;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int t = acc_row(rb, v, lane);
        if (t < r) wg[(int64_t)t * A.ld] = acc[rb][v];
      }
  }
}

// Wave-uniform copy of a descriptor read from global memory: the compiler treats values loaded
// from writable global memory as per-lane, which would turn every address derived from them into
// vector arithmetic; every 32-bit word is moved to a scalar register instead.
template <class T>
__device__ __forceinline__ T uniform_load(const T *p) {
  static_assert(sizeof(T) % 4 == 0, "descriptor size");
  T v = *p;
  unsigned w[sizeof(T) / 4];
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (unsigned i = 0; i < sizeof(T) / 4; ++i) w[i] = __builtin_amdgcn_readfirstlane(w[i]);
  __builtin_memcpy(&v, w, sizeof(T));
  return v;
}

// -------------------------------------------------------------------------------------------
// Chain kernel: persistent workgroups drain a queue of chunks (runs of tiles of one program); each
// chunk writes its own partial slot, so the result does not depend on which workgroup took it.

// S[t][e] = src(t, e) for t < r, e < w, with four global loads in flight per thread
template <int NT, class Src>
__device__ __forceinline__ void load_slab(double *S, int ldn, int r, int w, int tid, Src src) {
  const int n = r * w;
  for (int x0 = 0; x0 < n; x0 += 4 * NT) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int x = x0 + u * NT + tid;
      v[u] = x < n ? src(x / w, x % w) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int x = x0 + u * NT + tid;
      if (x < n) S[(size_t)(x / w) * ldn + x % w] = v[u];
    }
  }
}

// columns [0, w) of every node row copied to the columns of items 1 .. nit - 1
template <int NT>
__device__ __forceinline__ void replicate_items(double *S, int ldn, int r, int w, int nit, int tid) {
  const int m = (nit - 1) * w;
  for (int x = tid; x < r * m; x += NT) {
    const int t = x / m, c = x % m;
    S[(size_t)t * ldn + w + c] = S[(size_t)t * ldn + c % w];
  }
}

// MB: largest block of the launch's programs (register arrays of operator and table rows).
// The U factor sits in LDS except in the 64-column, four-wave tile, which leaves it in global
// memory so that two workgroups fit one CU (LDS and registers): while one waits at a barrier, the
// other's waves issue.
template <int RB, int NW, int NTILE>
constexpr bool chain_u_in_lds() { return !(NW == 4 && NTILE == 64); }

template <int RB, int NW, int NTILE, int MB>
__global__ void __launch_bounds__(64 * NW, (chain_u_in_lds<RB, NW, NTILE>() ? 1 : 2)) chain_kernel(ChainBatch B) {
  constexpr int RP = 16 * RB, NT = 64 * NW, LDN = NTILE + 4;
  constexpr int NCBT = NTILE / 16;           // 16-column blocks in a tile
  constexpr int NCB = NCBT * RB / NW;        // column blocks per wave in the row-partitioned GEMMs
  static_assert(NW % RB == 0 && (NCBT * RB) % NW == 0 && NCB >= 1, "waves must tile the (row, column) blocks");
  constexpr bool PF = MB <= 6;               // table rows read a round ahead (register budget)
  constexpr int VU = 2;                      // vertex rows per pass
  extern __shared__ double lds[];
  __shared__ int qnext;
  double *S0 = lds, *S1 = lds + RP * LDN;
  constexpr bool ULDS = chain_u_in_lds<RB, NW, NTILE>();
  double *Ul = lds + 2 * RP * LDN, *Dl = Ul + RP * (RP - 1) / 2;
  double *ipref = ULDS ? Dl + RP * 4 : Ul;                     // per item prefactor (NTILE)
  unsigned long long *imax = (unsigned long long *)(ipref + NTILE);  // per item max |T|
  int *ipole = (int *)(imax + NTILE);                          // (NTILE, kMaxLines)
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int rbw = wv % RB, grp = wv / RB;  // this wave's row block and column group
  constexpr int NGRP = NW / RB;              // column groups
  const double *cur = nullptr;               // DLR fragments of the problem whose U sits in LDS
  double *cache = B.scratch + (size_t)blockIdx.x * B.scratch_stride;  // this workgroup's shared-prefix T
  int cache_seg = -1, cache_prog = -1, cache_key = -1, cache_rows = 0, cache_cols = 0;
  for (;;) {
  __syncthreads();
  if (tid == 0) qnext = atomicAdd(B.queue, 1);
  __syncthreads();
  const int gchunk = __builtin_amdgcn_readfirstlane(qnext);  // wave-uniform: descriptors load through the scalar cache
  if (gchunk >= B.total) break;
  const int sidx = __builtin_amdgcn_readfirstlane(find_segment(B.prefix, B.nsegs, gchunk));
  const ChainArgs A = uniform_load(B.segs + sidx);
  const DlrFrag &D = A.D;
  const int r = A.r;
  const int nk = (r + 3) >> 2;
  if (D.LiP != cur) {
    // another problem: clean slabs (the padding rows start at its own rank) and its U factor
    for (int x = tid; x < RP * LDN; x += NT) { S0[x] = 0.0; S1[x] = 0.0; }
    if (ULDS) load_u<RB>(D, Ul, Dl, tid, NT);
    cur = D.LiP;
    __syncthreads();
  }
  const DChunk ch = uniform_load(A.chunks + __builtin_amdgcn_readfirstlane(gchunk - B.prefix[sidx]));
  const DProg P = uniform_load(A.progs + ch.prog);
  const int64_t slot = ch.slot;
  bool first_tile = true;
  for (int ti = ch.t0; ti < ch.t1; ++ti) {
    const int tstart = ti * P.items_per_tile;
    const int P_items = min(P.items_per_tile, P.ntup - tstart);
    // Items whose first-line pole agrees share the ops [0, split): those run on item 0 only, or
    // are read back from this workgroup's cache of the last shared prefix, and the result is
    // copied to the other items. Pole tuples run with the first line slowest.
    const int pkey = P.split > 0 ? tstart / P.share_len : -1;
    const bool shared = P.split > 0 && pkey == (tstart + P_items - 1) / P.share_len;
    const bool hit = shared && cache_seg == sidx && cache_prog == ch.prog && cache_key == pkey;
    __syncthreads();
    // pole tuple of each item, prefactor
    for (int it = tid; it < P_items; it += NT) {
      int u = P.tup_off >= 0 ? A.tuples[P.tup_off + tstart + it] : tstart + it;
      double pref = 1.0;
      int ls[kMaxLines];
#pragma unroll
      for (int x = kMaxLines - 1; x >= 0; --x) {
        ls[x] = 0;
        if (x < P.nlines) {
          ls[x] = A.plist[P.set_off[x] + u % P.set_len[x]];
          u /= P.set_len[x];
        }
        ipole[it * kMaxLines + x] = ls[x];
      }
#pragma unroll
      for (int x = 0; x < kMaxLines; ++x)
        if (x < P.nlines && P.pk_exp[x] != 0)
          pref *= A.kpow[((size_t)(P.pk_sign[x] > 0 ? 0 : 1) * A.maxexp + (P.pk_exp[x] - 1)) * A.p + ls[x]];
      ipref[it] = pref * (double)P.sign;
      imax[it] = 0ull;
    }
    int nit = shared ? 1 : P_items;  // items the next ops run on
    int rows = P.b1, cols = P.b1, oi0 = 0;
    if (hit) {
      rows = cache_rows;
      cols = cache_cols;
      oi0 = P.split;
      const int w = rows * cols;
      load_slab<NT>(S0, LDN, r, w, tid, [&](int t, int e) { return cache[(size_t)t * w + e]; });
    } else {  // T = G_b1, or its transpose: read once into item 0's columns, then copied to the others
      const int b1 = P.b1, w = b1 * b1, gtr = P.gtr;
      const double *Gb = A.G + P.g_init;
      load_slab<NT>(S0, LDN, r, w, tid, [&](int t, int e) { return Gb[(size_t)t * w + (gtr ? (e % b1) * b1 + e / b1 : e)]; });
      if (nit > 1) {
        __syncthreads();
        replicate_items<NT>(S0, LDN, r, w, nit, tid);
      }
    }
    __syncthreads();
    double *Scur = S0, *Soth = S1;
    for (int oi = oi0; oi < P.nops; ++oi) {
      if (shared && oi == P.split) {
        const int w = rows * cols;
        if (!hit) {
          for (int x = tid; x < r * w; x += NT) cache[x] = Scur[(size_t)(x / w) * LDN + x % w];
          cache_seg = sidx;
          cache_prog = ch.prog;
          cache_key = pkey;
          cache_rows = rows;
          cache_cols = cols;
        }
        replicate_items<NT>(Scur, LDN, r, w, P_items, tid);
        nit = P_items;
        __syncthreads();
      }
      const DOp op = uniform_load(A.ops + P.op0 + oi);
      if (op.type == 0) {  // vertex: T' = K F T
        const int bin = op.bin, bout = op.bout, win = bin * cols, wout = bout * cols;
        // one output row (t, item, i) per thread, VU rows per pass with their global reads issued
        // first: the operator row is loaded once and applied to every column; lanes of one
        // (t, item) read the same T entries (LDS broadcast)
        const int n = r * nit * bout, nib = nit * bout;
        const double *kt = op.ks > 0 ? A.kp : A.km;
        for (int x0 = tid; x0 < n; x0 += VU * NT) {
          double fr[VU][MB], kf[VU];
          int tt[VU], itt[VU], ii[VU];
#pragma unroll
          for (int u = 0; u < VU; ++u) {
            const int x = x0 + u * NT;
            const bool act = x < n;
            const int t = act ? x / nib : 0, rem = act ? x % nib : 0, it = rem / bout, i = rem % bout;
            tt[u] = t;
            itt[u] = it;
            ii[u] = i;
            const double *F = A.opmats + op.fbase + (op.bar_line >= 0 ? (int64_t)ipole[it * kMaxLines + op.bar_line] * op.fstride_l : 0);
#pragma unroll
            for (int k = 0; k < MB; ++k) fr[u][k] = (act && k < bin) ? F[i * bin + k] : 0.0;
            kf[u] = (act && op.ks != 0) ? kt[(size_t)ipole[it * kMaxLines + op.kx] * r + t] : 1.0;
          }
#pragma unroll
          for (int u = 0; u < VU; ++u) {
            if (x0 + u * NT >= n) break;
            const double *Tin = Scur + (size_t)tt[u] * LDN + itt[u] * win;
            double *Tout = Soth + (size_t)tt[u] * LDN + itt[u] * wout + ii[u] * cols;
            for (int j0 = 0; j0 < cols; j0 += MB) {  // MB independent column sums at a time
              double s[MB];
#pragma unroll
              for (int j = 0; j < MB; ++j) s[j] = 0.0;
#pragma unroll
              for (int k = 0; k < MB; ++k)
                if (k < bin) {
#pragma unroll
                  for (int j = 0; j < MB; ++j)
                    if (j0 + j < cols) s[j] += fr[u][k] * Tin[k * cols + j0 + j];
                }
#pragma unroll
              for (int j = 0; j < MB; ++j)
                if (j0 + j < cols) Tout[j0 + j] = kf[u] * s[j];
            }
          }
        }
        rows = bout;
        double *tmp = Scur; Scur = Soth; Soth = tmp;
        __syncthreads();
      } else if (op.type == 1) {  // edge integration
        const int b = op.bin, w = b * cols;
        // only the column blocks this tile's items occupy: group g of the waves takes blocks
        // [g cpg, min((g + 1) cpg, ncb)) of every row block
        const int ncb = (nit * w + 15) >> 4;
        const int cpg = (ncb + NGRP - 1) / NGRP;
        const int cbw = grp * cpg, ncw = min(cpg, ncb - cbw);
        // node-local phase layout (below): one output row (t, item, i) per thread, rounds of gpr
        // whole (t, item) groups; the groups past the last full round are split into single
        // elements when they fit one round
        const int ngroups = r * nit, gpr = NT / b;
        const int nfull = ngroups <= gpr ? ngroups : (ngroups / gpr) * gpr, ntail = ngroups - nfull;
        const bool split_tail = ntail > 0 && ntail * b * cols <= NT;
        const int gend = split_tail ? nfull : ngroups;
        const int gl = tid / b, irow = tid % b;
        // table rows f_i(t), (W f)_i(t) of group gi = (t, item) and row i (zeros when not act)
        auto load_elem = [&](bool act, int gi, int i, double(&f_)[MB], double(&wf_)[MB]) {
          const int t = act ? gi / nit : 0, it = act ? gi % nit : 0;
          int64_t e = op.ebase;
#pragma unroll
          for (int q = 0; q < kMaxLines; ++q)
            if (q < op.nlines) e += (int64_t)ipole[it * kMaxLines + op.lines[q]] * op.estride[q];
          const double *f = A.ftab + e + (size_t)t * b * b + i * b;
          const double *wf = f + (size_t)r * b * b;
#pragma unroll
          for (int k = 0; k < MB; ++k) {
            f_[k] = (act && k < b) ? f[k] : 0.0;
            wf_[k] = (act && k < b) ? wf[k] : 0.0;
          }
        };
        // this thread's rows in the round starting at group g0
        auto load_rows = [&](int g0, double(&f_)[MB], double(&wf_)[MB]) {
          load_elem(gl < gpr && g0 + gl < gend, g0 + gl, irow, f_, wf_);
        };
        // with small blocks the first round's rows are read now, so that their latency overlaps the
        // GEMMs, and each round reads the next round's rows before it computes (registers allow it)
        double fr[MB], wfr[MB];
        if (PF) load_rows(0, fr, wfr);
        {  // y = (L^{-1} P) T -> Soth
          v4d acc[NCB];
          zero_acc(acc);
          gemm_rows<NCB>(D.LiP, 4 * RB, rbw, 0, nk, Scur, LDN, cbw, ncw, acc, lane);
          store_rows<NCB>(acc, Soth, LDN, rbw, cbw, ncw, lane, r);
        }
        __syncthreads();
        for (int cb = wv; cb < ncb; cb += NW) {  // g = U^{-1} y in place, one column block per wave at a time
          v4d acc[RB];
          load_acc<RB>(acc, Soth, LDN, 16 * cb, lane);
          if (ULDS) usolve_lds<RB>(acc, Ul, Dl, r, lane);
          else usolve<RB>(acc, D, lane);
          store_acc<RB>(acc, Soth, LDN, 16 * cb, lane, r);
        }
        __syncthreads();
        {  // W g -> Scur (T is dead)
          v4d acc[NCB];
          zero_acc(acc);
          gemm_rows<NCB>(D.W, 4 * RB, rbw, 0, nk, Soth, LDN, cbw, ncw, acc, lane);
          store_rows<NCB>(acc, Scur, LDN, rbw, cbw, ncw, lane, r);
        }
        __syncthreads();
        // c = -(Wf o g + f o Wg) -> Soth, d = f o g -> Scur, in place. One output row (t, item, i)
        // per thread with the table rows f_i, (W f)_i in registers; rounds hold whole (t, item)
        // groups and 4 columns, whose reads all happen before any of them is overwritten. The next
        // round's table rows (or the tail element's) are read before the current round computes.
        {
          const int bc = b * cols;
          const bool tail_act = split_tail && tid < ntail * bc;
          const int tgi = nfull + (tail_act ? tid / bc : 0), tie = tail_act ? (tid % bc) / cols : 0, tje = tail_act ? tid % cols : 0;
          for (int g0 = 0; g0 < gend; g0 += gpr) {
            const int gi = g0 + gl;
            const bool act = gl < gpr && gi < gend;
            const int t = act ? gi / nit : 0, it = act ? gi % nit : 0;
            double fn[MB], wfn[MB];
            if (!PF) load_rows(g0, fr, wfr);
            else if (g0 + gpr < gend) load_rows(g0 + gpr, fn, wfn);
            else if (split_tail) load_elem(tail_act, tgi, tie, fn, wfn);
            double *gcol = Soth + (size_t)t * LDN + it * w, *wgcol = Scur + (size_t)t * LDN + it * w;
            for (int j0 = 0; j0 < cols; j0 += 4) {
              double cj[4], dj[4];
#pragma unroll
              for (int jj = 0; jj < 4; ++jj) {
                double s1 = 0.0, s2 = 0.0;
                const int j = j0 + jj;
                if (act && j < cols) {
#pragma unroll
                  for (int k = 0; k < MB; ++k)
                    if (k < b) {
                      const double g = gcol[k * cols + j], wg = wgcol[k * cols + j];
                      s1 += wfr[k] * g + fr[k] * wg;
                      s2 += fr[k] * g;
                    }
                }
                cj[jj] = -s1;
                dj[jj] = s2;
              }
              __syncthreads();
              if (act) {
#pragma unroll
                for (int jj = 0; jj < 4; ++jj)
                  if (j0 + jj < cols) {
                    gcol[irow * cols + j0 + jj] = cj[jj];
                    wgcol[irow * cols + j0 + jj] = dj[jj];
                  }
              }
              __syncthreads();
            }
            if (PF) {
#pragma unroll
              for (int k = 0; k < MB; ++k) {
                fr[k] = fn[k];
                wfr[k] = wfn[k];
              }
            }
          }
          if (split_tail) {  // one element (group, row, column) per thread; its rows are in fr, wfr
            if (!PF || gend == 0) load_elem(tail_act, tgi, tie, fr, wfr);
            const int t = tgi / nit, it = tgi % nit;
            double s1 = 0.0, s2 = 0.0;
            if (tail_act) {
              const double *gcol = Soth + (size_t)t * LDN + it * w, *wgcol = Scur + (size_t)t * LDN + it * w;
#pragma unroll
              for (int k = 0; k < MB; ++k)
                if (k < b) {
                  const double g = gcol[k * cols + tje], wg = wgcol[k * cols + tje];
                  s1 += wfr[k] * g + fr[k] * wg;
                  s2 += fr[k] * g;
                }
            }
            __syncthreads();
            if (tail_act) {
              Soth[(size_t)t * LDN + it * w + tie * cols + tje] = -s1;
              Scur[(size_t)t * LDN + it * w + tie * cols + tje] = s2;
            }
            __syncthreads();
          }
        }
        // h = beta (cf2it c + tcf2it d) -> Soth, stored after every wave has read both slabs
        {
          v4d acc[NCB];
          zero_acc(acc);
          gemm_rows<NCB>(D.M2, 8 * RB, rbw, 0, nk, Soth, LDN, cbw, ncw, acc, lane);
          gemm_rows<NCB>(D.M2, 8 * RB, rbw, 4 * RB, nk, Scur, LDN, cbw, ncw, acc, lane);
          __syncthreads();
          store_rows<NCB>(acc, Soth, LDN, rbw, cbw, ncw, lane, r);
        }
        // keep the padding rows of the slab that held d zero for the next GEMM that reads it
        double *tmp = Scur; Scur = Soth; Soth = tmp;
        __syncthreads();
      } else {  // zero-vertex pair
        const int bv = op.bin, bvp = op.bout, b1 = op.b1, b0 = op.cols, qs = op.qs, win = bv * b1, wout = bvp * b0;
        // out(t)[i][j] = sum_mu sum_a Fm_mu[i][a] sum_c T(t)[a][c] Fh_mu(t)[c][j] with the kappa
        // sums Fh_mu(t) = sum_kk h(t)[mu][kk] F_kk tabulated per (set, direction, block) at fbase,
        // (r, qs, b1, b0); one column of Fh_mu(t) in registers, the mu operators read with
        // wave-uniform addresses
        const int R = r * nit * b0, whole = R <= NT ? R : (R / NT) * NT;
        for (int x = tid; x < whole; x += NT) {
          const int t = x / (nit * b0), rem = x % (nit * b0), it = rem / b0, j = rem % b0;
          const double *Tt = Scur + (size_t)t * LDN + it * win;
          const double *Fht = A.opmats + op.fbase + (size_t)t * qs * b1 * b0 + j;
          double out[MB], fh[MB];
#pragma unroll
          for (int i = 0; i < MB; ++i) out[i] = 0.0;
#pragma unroll
          for (int c = 0; c < MB; ++c) fh[c] = (c < b1) ? Fht[(size_t)c * b0] : 0.0;
          for (int mu = 0; mu < qs; ++mu) {
            double fn[MB];  // the next mu's column, read before this one is used
#pragma unroll
            for (int c = 0; c < MB; ++c) fn[c] = (mu + 1 < qs && c < b1) ? Fht[(size_t)((mu + 1) * b1 + c) * b0] : 0.0;
            const double *Fm = A.opmats + op.mubase + (int64_t)mu * bvp * bv;
            for (int a = 0; a < bv; ++a) {
              double y = 0.0;
#pragma unroll
              for (int c = 0; c < MB; ++c)
                if (c < b1) y += Tt[a * b1 + c] * fh[c];
#pragma unroll
              for (int i = 0; i < MB; ++i)
                if (i < bvp) out[i] += Fm[i * bv + a] * y;
            }
#pragma unroll
            for (int c = 0; c < MB; ++c) fh[c] = fn[c];
          }
#pragma unroll
          for (int i = 0; i < MB; ++i)
            if (i < bvp) Soth[(size_t)t * LDN + it * wout + i * b0 + j] = out[i];
        }
        // the columns past the last full round: one output element (t, item, i, j) per thread
        for (int e = tid; e < (R - whole) * bvp; e += NT) {
          const int x = whole + e / bvp, i = e % bvp;
          const int t = x / (nit * b0), rem = x % (nit * b0), it = rem / b0, j = rem % b0;
          const double *Tt = Scur + (size_t)t * LDN + it * win;
          const double *Fht = A.opmats + op.fbase + (size_t)t * qs * b1 * b0 + j;
          double o = 0.0, fh[MB];
#pragma unroll
          for (int c = 0; c < MB; ++c) fh[c] = (c < b1) ? Fht[(size_t)c * b0] : 0.0;
          for (int mu = 0; mu < qs; ++mu) {
            double fn[MB];
#pragma unroll
            for (int c = 0; c < MB; ++c) fn[c] = (mu + 1 < qs && c < b1) ? Fht[(size_t)((mu + 1) * b1 + c) * b0] : 0.0;
            const double *Fmi = A.opmats + op.mubase + (int64_t)mu * bvp * bv + i * bv;
            for (int a = 0; a < bv; ++a) {
              double y = 0.0;
#pragma unroll
              for (int c = 0; c < MB; ++c)
                if (c < b1) y += Tt[a * b1 + c] * fh[c];
              o += Fmi[a] * y;
            }
#pragma unroll
            for (int c = 0; c < MB; ++c) fh[c] = fn[c];
          }
          Soth[(size_t)t * LDN + it * wout + i * b0 + j] = o;
        }
        rows = bvp;
        cols = b0;
        double *tmp = Scur; Scur = Soth; Soth = tmp;
        __syncthreads();
      }
    }
    if (P.out_off >= 0) {  // store mode: every item's T with its prefactor, time index fastest
      const int w = rows * cols;
      double *o = A.store + P.out_off + (size_t)tstart * r * w;
      for (int x = tid; x < P_items * w * r; x += NT) {
        const int c = x / r, t = x % r;  // c = item * w + element
        o[x] = Scur[(size_t)t * LDN + c] * ipref[c / w];
      }
      continue;
    }
    // final: per item max |T| (with the prefactor) and accumulation into the chunk's partial slot
    const int b0 = P.b0, w = b0 * b0;
    for (int x = tid; x < P_items * r; x += NT) {
      const int it = x / r, t = x % r;
      double mx = 0.0;
      for (int c = 0; c < w; ++c) mx = fmax(mx, fabs(Scur[(size_t)t * LDN + it * w + c] * ipref[it]));
      atomicMax(&imax[it], (unsigned long long)__double_as_longlong(mx));
    }
    __syncthreads();
    double *out = A.partial + slot;
    for (int x = tid; x < r * w; x += NT) {
      const int t = x / w, c = x % w;
      double s = first_tile ? 0.0 : out[x];
      for (int it = 0; it < P_items; ++it)
        if (__longlong_as_double((long long)imax[it]) > 1e-16) s += Scur[(size_t)t * LDN + it * w + c] * ipref[it];
      out[x] = s;
    }
    first_tile = false;
    (void)rows;
  }
  }
  // The queue resets itself for the next launch: every workgroup has taken its last ticket before
  // it counts itself out, so the last one out may clear both counters.
  if (tid == 0) {
    __threadfence();
    if (atomicAdd(B.queue + 1, 1) == (int)gridDim.x - 1) {
      B.queue[0] = 0;
      B.queue[1] = 0;
    }
  }
}

// Sigma[block] += scale * sum over the block's chunks of the partial slots. Thread (e, g) of a
// 16 x 16 workgroup sums chunks g, g + 16, ... of element e; the 16 group sums are then added in a
// fixed order, so the result is deterministic.
__global__ void __launch_bounds__(256) reduce_kernel(const int64_t *__restrict__ slot, const int *__restrict__ red_chunks,
                                                     const I2 *__restrict__ red_range, const int64_t *__restrict__ sig_off,
                                                     const int *__restrict__ sig_size, const double *__restrict__ partial,
                                                     double *__restrict__ sigma, double scale) {
  __shared__ double part[16][17];
  const int blk = blockIdx.y, e = threadIdx.x & 15, g = threadIdx.x >> 4;
  const int x = blockIdx.x * 16 + e;
  const I2 rg = red_range[blk];
  const bool live = x < sig_size[blk];
  double s = 0.0;
  if (live && rg.y > rg.x) {  // the block's slots are contiguous, sig_size[blk] doubles each
    const double *base = partial + slot[red_chunks[rg.x]] + x;
    const int64_t sz = sig_size[blk];
#pragma unroll 4
    for (int c = g; c < rg.y - rg.x; c += 16) s += base[(int64_t)c * sz];
  }
  part[g][e] = s;
  __syncthreads();
  if (g == 0 && live && rg.y > rg.x) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) t += part[q][e];
    sigma[sig_off[blk] + x] += scale * t;
  }
}

// Sigma of each problem, written: sum over its plans in order of scale * (chunk sums of that plan,
// split over 16 groups and combined in a fixed order). blockIdx.y runs over (problem, block) pairs.
__global__ void __launch_bounds__(256) reduce_assign_kernel(RedBatch B) {
  __shared__ double part[16][17];
  const int pidx = find_segment(B.prefix, B.nprobs, blockIdx.y);
  const RedProb &Q = B.probs[pidx];
  const int blk = blockIdx.y - B.prefix[pidx];
  const int e = threadIdx.x & 15, g = threadIdx.x >> 4;
  const int x = blockIdx.x * 16 + e;
  const bool live = x < Q.sig_size[blk];
  double acc = 0.0;
  for (int q = 0; q < Q.nplans; ++q) {
    const RedPlan &R = Q.plans[q];
    const I2 rg = R.red_range[blk];
    double s = 0.0;
    if (live && rg.y > rg.x) {  // the block's slots are contiguous, sig_size[blk] doubles each
      const double *base = R.partial + R.slot[R.red_chunks[rg.x]] + x;
      const int64_t sz = Q.sig_size[blk];
#pragma unroll 4
      for (int c = g; c < rg.y - rg.x; c += 16) s += base[(int64_t)c * sz];
    }
    __syncthreads();
    part[g][e] = s;
    __syncthreads();
    if (g == 0) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < 16; ++k) t += part[k][e];
      acc += R.scale * t;
    }
  }
  if (g == 0 && live) Q.sigma[Q.sig_off[blk] + x] = acc;
}

// chunk counter and exit counter of a persistent launch, cleared once when the queue is created
__global__ void reset_queue_kernel(int *q) {
  if (threadIdx.x < 2) q[threadIdx.x] = 0;
}

// One-time correlator contraction. Workgroup = chunk (pole tuples [k0, k1) of one group), wave = one
// output run at a time, lane = time node t. Per term and tuple a lane adds
//   Tr[L O_mu R O_kap](t) = sum_i sum_a (O_mu R)_ia (L^T O_kap^T)_ia
// with the operators as sparse rows (wave-uniform, scalar loads) and R, L^T read with the time index
// fastest (coalesced). Runs are summed over the chunk's tuples in a fixed order, so the result does
// not depend on the schedule; every chunk writes its whole (r, nout) slot.
constexpr int kCorrStageInts = 1024, kCorrStageVals = 512;  // per wave: a term's operators in LDS

template <bool LARGE>
__global__ void __launch_bounds__(256) corr_kernel(CorrArgs A) {
  __shared__ int s_cidx[LARGE ? 4 : 1][LARGE ? kCorrStageInts : 1];
  __shared__ double s_cval[LARGE ? 4 : 1][LARGE ? kCorrStageVals : 1];
  const CChunk ch = uniform_load(A.chunks + blockIdx.x);
  const int r = A.r, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, t = lane;
  double *slot = A.partial + (size_t)blockIdx.x * r * A.nout;
  for (int x = tid; x < r * A.nout; x += blockDim.x) slot[x] = 0.0;
  __syncthreads();
  for (int ru = ch.run0 + wv; ru < ch.run1; ru += 4) {
    const CRun run = uniform_load(A.runs + ru);
    double sum = 0.0;
    for (int q = run.t0; q < run.t1; ++q) {
      const CTerm T = uniform_load(A.terms + q);
      const int d0 = T.d0, d1 = T.d1, db = T.db, dc = T.dc;
      const int *mp = A.cidx + T.mu_csr, *mc = mp + d1 + 1;
      const int *kp = A.cidx + T.kap_csr, *kc = kp + db + 1;
      const double *mv = A.cval + T.mu_val, *kv = A.cval + T.kap_val;
      bool staged = false, single = false;
      int *smp = nullptr, *skp = nullptr, *skj = nullptr;
      double *smv = nullptr, *skv = nullptr, *skw = nullptr;
      if (LARGE) {
        // copy the term's two sparse operators (row pointers, columns, values) into this wave's part
        // of LDS, so that the loops below read them without global round trips
        const int nmu = __builtin_amdgcn_readfirstlane(mp[d1]), nkap = __builtin_amdgcn_readfirstlane(kp[db]);
        const int ni = (d1 + 1 + nmu) + (db + 1 + nkap) + db, nv = nmu + nkap + db;
        staged = ni <= kCorrStageInts && nv <= kCorrStageVals;
        if (staged) {
          __builtin_amdgcn_wave_barrier();  // the previous term's reads of these slots are done
          smp = s_cidx[wv];
          skp = smp + d1 + 1 + nmu;
          skj = skp + db + 1 + nkap;
          smv = s_cval[wv];
          skv = smv + nmu;
          skw = skv + nkap;
          bool one = true;  // every row of both operators holds at most one entry
          for (int x = lane; x < d1 + 1 + nmu; x += 64) smp[x] = mp[x];
          for (int x = lane; x < db + 1 + nkap; x += 64) skp[x] = kp[x];
          for (int x = lane; x < nmu; x += 64) smv[x] = mv[x];
          for (int x = lane; x < nkap; x += 64) skv[x] = kv[x];
          for (int x = lane; x < d1; x += 64) one = one && mp[x + 1] - mp[x] <= 1;
          for (int x = lane; x < db; x += 64) {  // O_kap row x as (column, value), (0, 0) when empty
            const int f0 = kp[x], f1 = kp[x + 1];
            one = one && f1 - f0 <= 1;
            skj[x] = f1 > f0 ? kc[f0] : 0;
            skw[x] = f1 > f0 ? kv[f0] : 0.0;
          }
          single = __ballot(!one) == 0;
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
      }
      for (int k = ch.k0; k < ch.k1; ++k) {
        if (t >= r) break;
        const double *R = A.store + T.r_off + (size_t)k * d0 * db * r + t;  // element (x, a) at R[(x db + a) r]
        const double *L = A.store + T.l_off + (size_t)k * d1 * dc * r + t;  // element (i, j) at L[(i dc + j) r]
        if (LARGE) {  // the same sums in the same order, each element of the mu product in turn
          // operator rows from this wave's LDS copy when the term's operators fit it
          const int *cmp = staged ? smp : mp, *cmc = staged ? smp + d1 + 1 : mc;
          const int *ckp = staged ? skp : kp, *ckc = staged ? skp + db + 1 : kc;
          const double *cmv = staged ? smv : mv, *ckv = staged ? skv : kv;
          if (single) {  // the same products and sums with at most one entry per operator row: the
                         // row loads of R and L no longer wait on each other
            for (int i = 0; i < d1; ++i) {
              const int e0 = __builtin_amdgcn_readfirstlane(smp[i]);
              if (__builtin_amdgcn_readfirstlane(smp[i + 1]) == e0) continue;
              const double v = smv[e0];
              const double *Rx = R + (size_t)__builtin_amdgcn_readfirstlane(smp[d1 + 1 + e0]) * db * r;
              const double *Li = L + (size_t)i * dc * r;
#pragma unroll 8
              for (int a = 0; a < db; ++a) {
                const double nra = 0.0 + v * Rx[(size_t)a * r];
                const double qv = 0.0 + skw[a] * Li[(size_t)skj[a] * r];
                sum += nra * qv;
              }
            }
            continue;
          }
          for (int i = 0; i < d1; ++i) {
            const int e0 = __builtin_amdgcn_readfirstlane(cmp[i]), e1 = __builtin_amdgcn_readfirstlane(cmp[i + 1]);
            if (e0 == e1) continue;  // a zero row of O_mu adds nothing
            const double *Li = L + (size_t)i * dc * r;
            for (int a = 0; a < db; ++a) {
              double nra = 0.0;
              for (int e = e0; e < e1; ++e) nra += cmv[e] * R[((size_t)__builtin_amdgcn_readfirstlane(cmc[e]) * db + a) * r];
              double qv = 0.0;
              const int f0 = __builtin_amdgcn_readfirstlane(ckp[a]), f1 = __builtin_amdgcn_readfirstlane(ckp[a + 1]);
              for (int e = f0; e < f1; ++e) qv += ckv[e] * Li[(size_t)__builtin_amdgcn_readfirstlane(ckc[e]) * r];
              sum += nra * qv;
            }
          }
          continue;
        }
        for (int i = 0; i < d1; ++i) {
          double nr[kMaxB];
#pragma unroll
          for (int a = 0; a < kMaxB; ++a) nr[a] = 0.0;
          const int e0 = __builtin_amdgcn_readfirstlane(mp[i]), e1 = __builtin_amdgcn_readfirstlane(mp[i + 1]);
          for (int e = e0; e < e1; ++e) {
            const double v = mv[e];
            const double *Rx = R + (size_t)__builtin_amdgcn_readfirstlane(mc[e]) * db * r;
#pragma unroll
            for (int a = 0; a < kMaxB; ++a)
              if (a < db) nr[a] += v * Rx[(size_t)a * r];
          }
          const double *Li = L + (size_t)i * dc * r;
#pragma unroll
          for (int a = 0; a < kMaxB; ++a)
            if (a < db) {
              double qv = 0.0;
              const int f0 = __builtin_amdgcn_readfirstlane(kp[a]), f1 = __builtin_amdgcn_readfirstlane(kp[a + 1]);
              for (int e = f0; e < f1; ++e) qv += kv[e] * Li[(size_t)__builtin_amdgcn_readfirstlane(kc[e]) * r];
              sum += nr[a] * qv;
            }
        }
      }
    }
    if (t < r) slot[(size_t)t * A.nout + run.out] = sum;
  }
}

// out[e] (=, or += when accumulate) + sum over chunks of partial[c][e]: thread (e, g) of a 16 x 16
// workgroup sums chunks g, g + 16, ... of element e, and the 16 group sums are added in a fixed order.
__global__ void __launch_bounds__(256) corr_reduce_kernel(const double *__restrict__ partial, int nchunks, int n, double *out,
                                                          int accumulate) {
  __shared__ double part[16][17];
  const int e = threadIdx.x & 15, g = threadIdx.x >> 4;
  const int x = blockIdx.x * 16 + e;
  double s = 0.0;
  if (x < n)
    for (int c = g; c < nchunks; c += 16) s += partial[(size_t)c * n + x];
  part[g][e] = s;
  __syncthreads();
  if (g == 0 && x < n) {
    double acc = accumulate ? out[x] : 0.0, tt = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) tt += part[q][e];
    out[x] = acc + tt;
  }
}

// -------------------------------------------------------------------------------------------
// host-callable launchers

template <int RB, int NW, int NTILE>
static size_t lds_bytes() {
  constexpr int RP = 16 * RB, LDN = NTILE + 4;
  constexpr size_t u = chain_u_in_lds<RB, NW, NTILE>() ? (size_t)RP * (RP - 1) / 2 + RP * 4 : 0;
  return sizeof(double) * ((size_t)2 * RP * LDN + u + NTILE + NTILE) + sizeof(int) * NTILE * kMaxLines;
}

template <int RB, int NW>
static hipError_t launch_tab_t(const TabBatch &B, hipStream_t s) {
  constexpr int RP = 16 * RB, LDN = 16 * NW + 8;
  const size_t sh = sizeof(double) * ((size_t)2 * RP * LDN + RP * (RP - 1) / 2 + RP * 4);
  static bool ready = false;  // attribute set once, outside any stream capture after the first launch
  if (!ready) {
    hipError_t e = hipFuncSetAttribute((const void *)tab_kernel<RB, NW>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
    ready = true;
  }
  if (B.total <= 0) return hipSuccess;
  hipLaunchKernelGGL((tab_kernel<RB, NW>), dim3(B.total), dim3(64 * NW), sh, s, B);
  return hipGetLastError();
}

// workgroups of this instance resident on the device at once (set on first use, with the LDS
// attribute, outside any stream capture)
template <int RB, int NW, int NTILE, int MB>
static int &resident_of() {
  static int resident = 0;
  return resident;
}

template <int RB, int NW, int NTILE, int MB>
static hipError_t init_chain_t() {
  int &resident = resident_of<RB, NW, NTILE, MB>();
  const size_t sh = lds_bytes<RB, NW, NTILE>();
  const void *fn = (const void *)chain_kernel<RB, NW, NTILE, MB>;
  if (resident == 0) {
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
    int dev = 0, ncu = 0, per = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, 64 * NW, sh)) != hipSuccess) return e;
    resident = std::max(1, ncu * std::max(per, 1));
  }
  return hipSuccess;
}

// Both block-size instances of a tile width; the workgroup count of the width is the larger of
// the two, so that per-workgroup scratch sized by it serves either.
template <int RB, int NW, int NTILE>
static int init_width_t() {
  if (init_chain_t<RB, NW, NTILE, 4>() != hipSuccess || init_chain_t<RB, NW, NTILE, kMaxB>() != hipSuccess) return 0;
  return std::max(resident_of<RB, NW, NTILE, 4>(), resident_of<RB, NW, NTILE, kMaxB>());
}

// persistent grid: as many workgroups as fit on the device at once, at most one per chunk. The
// instance with 4-element register rows runs launches whose blocks all have at most 4 states.
template <int RB, int NW, int NTILE>
static hipError_t launch_chain_t(const ChainBatch &B, hipStream_t s) {
  if (init_width_t<RB, NW, NTILE>() <= 0) return hipErrorInvalidConfiguration;
  const size_t sh = lds_bytes<RB, NW, NTILE>();
  if (B.total <= 0) return hipSuccess;
  if (B.maxb < 1 || B.maxb > kMaxB) return hipErrorInvalidValue;
  if (B.maxb <= 4) {
    const int grid = std::min(B.total, resident_of<RB, NW, NTILE, 4>());
    hipLaunchKernelGGL((chain_kernel<RB, NW, NTILE, 4>), dim3(grid), dim3(64 * NW), sh, s, B);
  } else {
    const int grid = std::min(B.total, resident_of<RB, NW, NTILE, kMaxB>());
    hipLaunchKernelGGL((chain_kernel<RB, NW, NTILE, kMaxB>), dim3(grid), dim3(64 * NW), sh, s, B);
  }
  return hipGetLastError();
}

}  // namespace xca::dev

// entry points used by the host orchestration (context.cpp)
namespace xca::dev {

// Tile variants (waves, columns) selectable per plan: 128 columns on 8 waves, 64 on 4, 32 on 4.
int num_tile_variants() { return 3; }
int tile_columns(int variant) {
  static const int cols[] = {128, 64, 32};
  return cols[variant];
}

hipError_t launch_tab(int RB, const TabBatch &B, hipStream_t s) {
  switch (RB) {
    case 2: return launch_tab_t<2, 8>(B, s);
    case 4: return launch_tab_t<4, 8>(B, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_chain(int RB, int variant, const ChainBatch &B, hipStream_t s) {
  if (RB == 2) {
    switch (variant) {
      case 0: return launch_chain_t<2, 8, 128>(B, s);
      case 1: return launch_chain_t<2, 4, 64>(B, s);
      case 2: return launch_chain_t<2, 4, 32>(B, s);
    }
  } else if (RB == 4) {
    switch (variant) {
      case 0: return launch_chain_t<4, 8, 128>(B, s);
      case 1: return launch_chain_t<4, 4, 64>(B, s);
      case 2: return launch_chain_t<4, 4, 32>(B, s);
    }
  }
  return hipErrorInvalidValue;
}

int chain_resident(int RB, int variant) {
  if (RB == 2) {
    switch (variant) {
      case 0: return init_width_t<2, 8, 128>();
      case 1: return init_width_t<2, 4, 64>();
      case 2: return init_width_t<2, 4, 32>();
    }
  } else if (RB == 4) {
    switch (variant) {
      case 0: return init_width_t<4, 8, 128>();
      case 1: return init_width_t<4, 4, 64>();
      case 2: return init_width_t<4, 4, 32>();
    }
  }
  return 0;
}

hipError_t launch_reduce(const int64_t *slot, const int *red_chunks, const I2 *red_range, const int64_t *sig_off, const int *sig_size,
                         int nblocks, int max_size, const double *partial, double *sigma, double scale, hipStream_t s) {
  dim3 grid((max_size + 15) / 16, nblocks);
  hipLaunchKernelGGL(reduce_kernel, grid, dim3(256), 0, s, slot, red_chunks, red_range, sig_off, sig_size, partial, sigma, scale);
  return hipGetLastError();
}

hipError_t launch_reduce_assign(const RedBatch &B, int max_size, hipStream_t s) {
  if (B.total <= 0) return hipSuccess;
  dim3 grid((max_size + 15) / 16, B.total);
  hipLaunchKernelGGL(reduce_assign_kernel, grid, dim3(256), 0, s, B);
  return hipGetLastError();
}

size_t chain_args_size() { return sizeof(ChainArgs); }

hipError_t init_queue(int *q, hipStream_t s) {
  hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(64), 0, s, q);
  return hipGetLastError();
}

template <int RB, int NW>
static hipError_t launch_refl_t(const ReflArgs &A, hipStream_t s) {
  constexpr int RP = 16 * RB, LDN = 16 * NW + 8;
  const size_t sh = sizeof(double) * ((size_t)2 * RP * LDN);
  static bool ready = false;
  if (!ready) {
    hipError_t e = hipFuncSetAttribute((const void *)refl_kernel<RB, NW>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
    ready = true;
  }
  if (A.njobs <= 0) return hipSuccess;
  hipLaunchKernelGGL((refl_kernel<RB, NW>), dim3(A.njobs), dim3(64 * NW), sh, s, A);
  return hipGetLastError();
}

int refl_columns() { return 16 * 8; }

hipError_t launch_refl(int RB, const ReflArgs &A, hipStream_t s) {
  return RB == 2 ? launch_refl_t<2, 8>(A, s) : launch_refl_t<4, 8>(A, s);
}

template <int RB, int NW>
static hipError_t launch_wv2c_t(const WV2C &A, hipStream_t s) {
  constexpr int RP = 16 * RB, LDN = 16 * NW + 8;
  const size_t sh = sizeof(double) * ((size_t)2 * RP * LDN);
  static bool ready = false;
  if (!ready) {
    hipError_t e = hipFuncSetAttribute((const void *)wv2c_kernel<RB, NW>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
    ready = true;
  }
  if (A.nitem <= 0 || A.ncol <= 0) return hipSuccess;
  const int cblocks = (A.ncol + 16 * NW - 1) / (16 * NW);
  hipLaunchKernelGGL((wv2c_kernel<RB, NW>), dim3(cblocks * A.nitem), dim3(64 * NW), sh, s, A);
  return hipGetLastError();
}

hipError_t launch_wv2c(int RB, const WV2C &A, hipStream_t s) {
  return RB == 2 ? launch_wv2c_t<2, 8>(A, s) : launch_wv2c_t<4, 8>(A, s);
}

// Set the launch attributes (dynamic LDS) of every kernel instance for this RB, and the resident
// workgroup counts of every chain variant, before any launch a stream capture may record. Each
// launcher sets its attribute on its first call and returns at once on an empty batch.
hipError_t prepare_kernels(int RB, hipStream_t s) {
  hipError_t e;
  if ((e = launch_tab(RB, TabBatch{}, s)) != hipSuccess) return e;
  if ((e = launch_refl(RB, ReflArgs{}, s)) != hipSuccess) return e;
  if ((e = launch_wv2c(RB, WV2C{}, s)) != hipSuccess) return e;
  for (int v = 0; v < num_tile_variants(); ++v)
    if (chain_resident(RB, v) <= 0) return hipErrorInvalidConfiguration;
  return hipSuccess;
}

hipError_t launch_corr(const CorrArgs &A, hipStream_t s) {
  if (A.nchunks <= 0) return hipSuccess;
  if (A.large) hipLaunchKernelGGL(corr_kernel<true>, dim3(A.nchunks), dim3(256), 0, s, A);
  else hipLaunchKernelGGL(corr_kernel<false>, dim3(A.nchunks), dim3(256), 0, s, A);
  return hipGetLastError();
}

hipError_t launch_corr_reduce(const double *partial, int nchunks, int n, double *out, int accumulate, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(corr_reduce_kernel, dim3((n + 15) / 16), dim3(256), 0, s, partial, nchunks, n, out, accumulate);
  return hipGetLastError();
}

}  // namespace xca::dev

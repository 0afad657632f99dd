// Kernels of the large-block Σ engine: a batched FP64 MFMA GEMM with per-item operand offsets and a
// per-(item, node) scale, the fused node-local products of an edge, edge-table inputs, the
// zero-vertex operator sums, and the per-item skip test and Σ accumulation. The node-space part of an
// edge (vals2coefs and W) runs in engine.hip's wv2c_kernel.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "wide.hpp"

namespace xca::dev {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

}  // namespace

// 64 x 64 output tile per workgroup, 4 waves of 32 x 32 (2 x 2 MFMA blocks of 16 x 16), K staged
// through LDS 32 at a time with the next K tile prefetched into registers while the current one is
// multiplied. MFMA f64 16x16x4 lane layout: A[i][k] in lane i + 16k, B[k][j] in lane j + 16k,
// D(i, j) in lane j + 16 (i mod 4), register i / 4.
__global__ void __launch_bounds__(256) wgemm_kernel(WGemm G) {
  constexpr int TM = 64, TN = 64, TK = 32, PER = TM * TK / 256;  // 8 A and 8 B elements per thread
  __shared__ double As[TK][TM + 1];
  __shared__ double Bs[TK][TN + 1];
  const int tilesN = (G.N + TN - 1) / TN;
  const int tm = blockIdx.x / tilesN, tn = blockIdx.x % tilesN;
  const int q = blockIdx.y, u = q / G.nt, t = q % G.nt;
  const double *A = G.A + (G.a_off ? G.a_off[u] : 0) + (int64_t)u * G.a_si + (int64_t)t * G.a_st;
  const double *B = G.B + (int64_t)u * G.b_si + (int64_t)t * G.b_st;
  double *C = G.C + (int64_t)u * G.c_si + (int64_t)t * G.c_st;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m0 = tm * TM, n0 = tn * TN;
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
  v4d acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = v4d{0.0, 0.0, 0.0, 0.0};
  double ra[PER], rb[PER];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int x = tid + 256 * e;
      const int i = x / TK, kk = x % TK, gi = m0 + i, gk = k0 + kk;
      ra[e] = (gi < G.M && gk < G.K) ? A[(int64_t)gi * G.lda + gk] : 0.0;
      const int kb = x / TN, j = x % TN, gkb = k0 + kb, gj = n0 + j;
      rb[e] = (gkb < G.K && gj < G.N) ? B[(int64_t)gkb * G.ldb + gj] : 0.0;
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < G.K; k0 += TK) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int x = tid + 256 * e;
      As[x % TK][x / TK] = ra[e];
      Bs[x / TN][x % TN] = rb[e];
    }
    __syncthreads();
    if (k0 + TK < G.K) fetch(k0 + TK);
    const int kend = min(TK, G.K - k0);
#pragma unroll
    for (int kk = 0; kk < TK; kk += 4) {
      if (kk < kend) {
        const int kr = kk + (lane >> 4), c = lane & 15;
        const double a0 = As[kr][wm + c], a1 = As[kr][wm + 16 + c];
        const double b0 = Bs[kr][wn + c], b1 = Bs[kr][wn + 16 + c];
        acc[0][0] = mfma(a0, b0, acc[0][0]);
        acc[0][1] = mfma(a0, b1, acc[0][1]);
        acc[1][0] = mfma(a1, b0, acc[1][0]);
        acc[1][1] = mfma(a1, b1, acc[1][1]);
      }
    }
    __syncthreads();
  }
  const double s = G.S ? G.S[G.s_off[u] + t] : 1.0;
  const double al = G.alpha * s;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = m0 + wm + 16 * bi + 4 * v + (lane >> 4), j = n0 + wn + 16 * bj + (lane & 15);
        if (i < G.M && j < G.N) {
          double *cp = C + (int64_t)i * G.ldc + j;
          const double val = al * acc[bi][bj][v];
          *cp = (G.beta != 0.0) ? val + G.beta * *cp : val;
        }
      }
}

// dst[u][x] = src[x] for x < n
__global__ void winit_kernel(const double *__restrict__ src, double *__restrict__ dst, int64_t n, int nitem, int ld) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n * nitem) return;
  int64_t x = g % n;
  if (ld > 0) {  // node block t, element (i, j) from (j, i)
    const int64_t bb = (int64_t)ld * ld, e = x % bb;
    x = x - e + (e % ld) * ld + e / ld;
  }
  dst[g] = src[x];
}

// Correlator store: item u's T (r x n, node-major) to dst with the node index fastest, times pref[u]
__global__ void wstore_kernel(const double *__restrict__ src, double *__restrict__ dst, int64_t n, int r, int nitem,
                              const double *__restrict__ pref) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = n * r;
  if (g >= per * nitem) return;
  const int64_t u = g / per, x = g % per;
  const int64_t e = x / r;
  const int t = (int)(x % r);
  dst[g] = pref[u] * src[u * per + (int64_t)t * n + e];
}

// Edge-table inputs: dst[e] = G (r x n) times the K factors of the table's lines for entry e
// (line q's pole = (e / p^q) mod p, sign ls[q])
__global__ void wgk_kernel(const double *__restrict__ G, double *__restrict__ dst, int r, int64_t n, int nent, int nlines, int ls0, int ls1,
                           int ls2, int p, const double *__restrict__ kp, const double *__restrict__ km, int ld) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)r * n;
  if (g >= per * nent) return;
  const int e = (int)(g / per);
  const int64_t x = g % per;
  const int t = (int)(x / n);
  int64_t xs = x;
  if (ld > 0) {  // element (i, j) of the transposed block
    const int64_t e = x % n;
    xs = x - e + (e % ld) * ld + e / ld;
  }
  double v = G[xs];
  const int ls[3] = {ls0, ls1, ls2};
  int ee = e;
  for (int q = 0; q < nlines; ++q) {
    const int l = ee % p;
    ee /= p;
    v *= (ls[q] > 0 ? kp : km)[(int64_t)l * r + t];
  }
  dst[g] = v;
}

// Zero-vertex operator sums, side by side per node: dst[t] is b1 x (qs b0) with columns mu b0 + j,
// Fh_mu(t) = sum_kap h(t, mu, kap) F_kap, h = Delta(t)[mu][kap] for a forward line,
// -Delta_refl(t)[kap][mu] otherwise (block_sparse_backbone.cpp:216-221)
__global__ void wfh_kernel(double *__restrict__ dst, const double *__restrict__ F, int qs, int b1, int b0, int r, int nflav, int forward,
                           int o0, int o1, int o2, int o3, int o4, int o5, int o6, int o7, const double *__restrict__ delta,
                           const double *__restrict__ delta_refl) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t w = (int64_t)b1 * b0;
  if (g >= (int64_t)qs * r * w) return;
  const int mu = (int)(g / (r * w));
  const int t = (int)((g / w) % r);
  const int64_t x = g % w;
  const int c = (int)(x / b0), j = (int)(x % b0);
  const int orb[8] = {o0, o1, o2, o3, o4, o5, o6, o7};
  double s = 0.0;
  for (int k = 0; k < qs; ++k) {
    const double h = forward ? delta[((int64_t)t * nflav + orb[mu]) * nflav + orb[k]] : -delta_refl[((int64_t)t * nflav + orb[k]) * nflav + orb[mu]];
    s += h * F[(int64_t)k * w + x];
  }
  dst[(int64_t)t * qs * w + (int64_t)c * qs * b0 + (int64_t)mu * b0 + j] = s;
}

// vmax[u] = max(vmax[u], max |src[u][x]|): slices of one item per workgroup (blockIdx.y = item),
// combined with an unsigned max on the bit patterns, which orders non-negative doubles
__global__ void __launch_bounds__(256) wmax_kernel(const double *__restrict__ src, int64_t n, double *vmax) {
  __shared__ double red[256];
  const int u = blockIdx.y;
  double m = 0.0;
  for (int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x; x < n; x += (int64_t)gridDim.x * 256) m = fmax(m, fabs(src[(int64_t)u * n + x]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax((unsigned long long *)(vmax + u), (unsigned long long)__double_as_longlong(red[0]));
}

__global__ void wzero_kernel(double *dst, int64_t n) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) dst[g] = 0.0;
}

// Node-local products of an edge, fused: one 64 x 64 tile of C[u][k] and D[u][k] per workgroup,
// f / Wf tiles and g / Wg tiles staged together through LDS, 32 of K at a time, the next K tile
// prefetched into registers.
__global__ void __launch_bounds__(256) wlocal_kernel(WLocal L) {
  constexpr int TM = 64, TN = 64, TK = 32, PER = TM * TK / 256;
  __shared__ double Af[TK][TM + 1], Aw[TK][TM + 1], Bg[TK][TN + 1], Bw[TK][TN + 1];
  const int b = L.b, cols = L.cols;
  const int tilesN = (cols + TN - 1) / TN;
  const int tm = blockIdx.x / tilesN, tn = blockIdx.x % tilesN;
  const int q = blockIdx.y, u = q / L.r, k = q % L.r;
  const double *f = L.F + L.f_off[u] + (int64_t)k * b * b, *wf = f + L.wf_shift;
  const int64_t gofs = (int64_t)u * L.g_si + (int64_t)k * b * cols;
  const double *g = L.Gk + gofs, *wg = L.WGk + gofs;
  const int64_t cofs = (int64_t)u * L.c_si + (int64_t)k * b * cols;
  double *C = L.C + cofs, *D = L.D + cofs;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m0 = tm * TM, n0 = tn * TN;
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
  v4d cc[2][2], dd[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) cc[bi][bj] = dd[bi][bj] = v4d{0.0, 0.0, 0.0, 0.0};
  double rf[PER], rw[PER], rg[PER], rh[PER];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int x = tid + 256 * e;
      const int i = x / TK, kk = x % TK, gi = m0 + i, gk = k0 + kk;
      const bool oka = gi < b && gk < b;
      rf[e] = oka ? f[(int64_t)gi * b + gk] : 0.0;
      rw[e] = oka ? wf[(int64_t)gi * b + gk] : 0.0;
      const int kb = x / TN, j = x % TN, gkb = k0 + kb, gj = n0 + j;
      const bool okb = gkb < b && gj < cols;
      rg[e] = okb ? g[(int64_t)gkb * cols + gj] : 0.0;
      rh[e] = okb ? wg[(int64_t)gkb * cols + gj] : 0.0;
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < b; k0 += TK) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int x = tid + 256 * e;
      Af[x % TK][x / TK] = rf[e];
      Aw[x % TK][x / TK] = rw[e];
      Bg[x / TN][x % TN] = rg[e];
      Bw[x / TN][x % TN] = rh[e];
    }
    __syncthreads();
    if (k0 + TK < b) fetch(k0 + TK);
    const int kend = min(TK, b - k0);
#pragma unroll
    for (int kk = 0; kk < TK; kk += 4) {
      if (kk < kend) {
        const int kr = kk + (lane >> 4), c = lane & 15;
        const double f0 = Af[kr][wm + c], f1 = Af[kr][wm + 16 + c];
        const double w0 = Aw[kr][wm + c], w1 = Aw[kr][wm + 16 + c];
        const double g0 = Bg[kr][wn + c], g1 = Bg[kr][wn + 16 + c];
        const double h0 = Bw[kr][wn + c], h1 = Bw[kr][wn + 16 + c];
        cc[0][0] = mfma(w0, g0, mfma(f0, h0, cc[0][0]));
        cc[0][1] = mfma(w0, g1, mfma(f0, h1, cc[0][1]));
        cc[1][0] = mfma(w1, g0, mfma(f1, h0, cc[1][0]));
        cc[1][1] = mfma(w1, g1, mfma(f1, h1, cc[1][1]));
        dd[0][0] = mfma(f0, g0, dd[0][0]);
        dd[0][1] = mfma(f0, g1, dd[0][1]);
        dd[1][0] = mfma(f1, g0, dd[1][0]);
        dd[1][1] = mfma(f1, g1, dd[1][1]);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = m0 + wm + 16 * bi + 4 * v + (lane >> 4), j = n0 + wn + 16 * bj + (lane & 15);
        if (i < b && j < cols) {
          C[(int64_t)i * cols + j] = -cc[bi][bj][v];
          D[(int64_t)i * cols + j] = dd[bi][bj][v];
        }
      }
}

// Sigma block += scale * sum over items, in order, of the kept items' pref * T
__global__ void wfinal_kernel(const double *__restrict__ src, int64_t n, int nitem, const double *__restrict__ vmax,
                              const double *__restrict__ pref, double *dst, double scale) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n) return;
  double s = 0.0;
  for (int u = 0; u < nitem; ++u)
    if (vmax[u] * fabs(pref[u]) > 1e-16) s += pref[u] * src[(int64_t)u * n + x];
  dst[x] += scale * s;
}

hipError_t launch_wide(const WLaunch &L, const WArgs &A, double scale, hipStream_t s) {
  switch (L.kind) {
    case W_GEMM: {
      const WGemm &G = L.g;
      if (G.nitem * G.nt <= 0 || G.M <= 0 || G.N <= 0) return hipSuccess;
      dim3 grid(((G.M + 63) / 64) * ((G.N + 63) / 64), G.nitem * G.nt);
      hipLaunchKernelGGL(wgemm_kernel, grid, dim3(256), 0, s, G);
      break;
    }
    case W_INIT: {
      const int64_t tot = L.n * L.nitem;
      hipLaunchKernelGGL(winit_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, L.src, L.dst, L.n, L.nitem, L.ld);
      break;
    }
    case W_V2C:
      return launch_wv2c(A.RB, L.v, s);
    case W_LOCAL: {
      const WLocal &W = L.loc;
      dim3 grid(((W.b + 63) / 64) * ((W.cols + 63) / 64), W.nitem * W.r);
      hipLaunchKernelGGL(wlocal_kernel, grid, dim3(256), 0, s, W);
      break;
    }
    case W_ZERO:
      hipLaunchKernelGGL(wzero_kernel, dim3((unsigned)((L.n + 255) / 256)), dim3(256), 0, s, L.dst, L.n);
      break;
    case W_GK: {
      const int64_t tot = (int64_t)A.r * L.n * L.nitem;
      hipLaunchKernelGGL(wgk_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, L.src, L.dst, A.r, L.n, L.nitem, L.nlines, L.ls[0],
                         L.ls[1], L.ls[2], L.p, A.kp, A.km, L.ld);
      break;
    }
    case W_FH: {
      const int64_t tot = (int64_t)L.qs * A.r * L.b1 * L.b0;
      hipLaunchKernelGGL(wfh_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, L.dst, L.src, L.qs, L.b1, L.b0, A.r, L.nflav,
                         L.forward, L.orb[0], L.orb[1], L.orb[2], L.orb[3], L.orb[4], L.orb[5], L.orb[6], L.orb[7], A.delta, A.delta_refl);
      break;
    }
    case W_MAX: {
      const int slices = (int)std::min<int64_t>(64, (L.n + 4095) / 4096);
      hipLaunchKernelGGL(wmax_kernel, dim3(slices, L.nitem), dim3(256), 0, s, L.src, L.n, L.vmax);
      break;
    }
    case W_FINAL:
      hipLaunchKernelGGL(wfinal_kernel, dim3((unsigned)((L.n + 255) / 256)), dim3(256), 0, s, L.src, L.n, L.nitem, L.vmax, L.pref, L.dst,
                         scale);
      break;
    case W_STORE: {
      const int64_t tot = (int64_t)A.r * L.n * L.nitem;
      hipLaunchKernelGGL(wstore_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, L.src, L.dst, L.n, A.r, L.nitem, L.pref);
      break;
    }
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace xca::dev

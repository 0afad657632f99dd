// Pseudo-particle Dyson equation and the iteration helpers on the device (FP64).
//
// Per Green's-function block b of dimension n (dyson_it_ppsc::solve, c++/triqs_xca/dlr_dyson_ppsc.hpp:135-168):
//   sigc = vals2coefs(Sigma_b)
//   ggs  = -(convolve_TO(beta, g0c, sigc) + eta G0_b)
//   S    = I + convmat(vals2coefs(ggs))          (r n x r n, acting on one column of G)
//   G_b  = S^{-1} G0_b  (pivoted LU, n right-hand sides), then hermitized (block_sparse_solver.py:579-587).
// vals2coefs is the LU solve with the DLR interpolation factors (substitution per column, backward
// stable). The convolution matrix is first assembled in coefficient space from the same time-ordered
// convolution formula as the self-energy kernels,
//   Mc[(t,i),(k',i')] = beta ( -cf2it(t,k') (W f)_k'[i,i'] + tcf2it(t,k') f_k'[i,i'] - sum_k cf2it(t,k) f_k[i,i'] W(k,k') ),
// and then composed with vals2coefs from the right by transposed LU solves, C^T = cf2it^{-T} Mc^T.
// Multiplying by an explicit inverse of cf2it instead amplifies rounding by its condition number and
// leaves the self-consistency loop stalled at ~1e-7.
#include <hip/hip_runtime.h>

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

// in-place LU solve of one column (stride between nodes) with the DLR interpolation factors
__device__ void lu_solve_col(const double *__restrict__ lu, const int *__restrict__ piv, int r, double *x, int stride) {
  for (int k = 0; k < r; ++k)
    if (piv[k] != k) {
      double a = x[(size_t)k * stride], b = x[(size_t)piv[k] * stride];
      x[(size_t)k * stride] = b;
      x[(size_t)piv[k] * stride] = a;
    }
  for (int i = 1; i < r; ++i) {
    double s = x[(size_t)i * stride];
    for (int k = 0; k < i; ++k) s -= lu[i * r + k] * x[(size_t)k * stride];
    x[(size_t)i * stride] = s;
  }
  for (int i = r - 1; i >= 0; --i) {
    double s = x[(size_t)i * stride];
    for (int k = i + 1; k < r; ++k) s -= lu[i * r + k] * x[(size_t)k * stride];
    x[(size_t)i * stride] = s / lu[i * r + i];
  }
}

// in-place solve of cf2it^T y = x for one column: U^T z = x, L^T w = z, then the row interchanges
// undone in reverse order
__device__ void lu_solve_t_col(const double *__restrict__ lu, const int *__restrict__ piv, int r, double *x, int stride) {
  for (int i = 0; i < r; ++i) {
    double s = x[(size_t)i * stride];
    for (int k = 0; k < i; ++k) s -= lu[k * r + i] * x[(size_t)k * stride];
    x[(size_t)i * stride] = s / lu[i * r + i];
  }
  for (int i = r - 2; i >= 0; --i) {
    double s = x[(size_t)i * stride];
    for (int k = i + 1; k < r; ++k) s -= lu[k * r + i] * x[(size_t)k * stride];
    x[(size_t)i * stride] = s;
  }
  for (int k = r - 1; k >= 0; --k)
    if (piv[k] != k) {
      double a = x[(size_t)k * stride], b = x[(size_t)piv[k] * stride];
      x[(size_t)k * stride] = b;
      x[(size_t)piv[k] * stride] = a;
    }
}

// Phase 1 (one workgroup per block): sigc, ggs, vals2coefs(ggs), W f, and the system matrix.
__global__ void __launch_bounds__(256) dyson_build_kernel(DysonBatch Bt) {
  const int pidx = find_seg(Bt.prefix, Bt.nprobs, blockIdx.x);
  const DysonArgs &A = Bt.probs[pidx];
  const int b = blockIdx.x - Bt.prefix[pidx], n = A.dims[b], r = A.r, nn = n * n, N = r * n;
  const double eta = *A.eta;
  const int tid = threadIdx.x, NT = blockDim.x;
  const size_t go = A.goff[b];
  const double *G0 = A.G0 + go, *g0c = A.g0c + go, *Wg0c = A.Wg0c + go;
  double *sigc = A.sigc + go, *ggs = A.ggs + go, *wf = A.wf + go;
  // sigc = vals2coefs(Sigma_b)
  for (int x = tid; x < r * nn; x += NT) sigc[x] = A.sigma[go + x];
  __syncthreads();
  for (int c = tid; c < nn; c += NT) lu_solve_col(A.lu, A.piv, r, sigc + c, nn);
  __syncthreads();
  // W sigc -> wf (scratch)
  for (int x = tid; x < r * nn; x += NT) {
    const int t = x / nn, c = x % nn;
    double s = 0.0;
    for (int k = 0; k < r; ++k) s += A.W[t * r + k] * sigc[(size_t)k * nn + c];
    wf[x] = s;
  }
  __syncthreads();
  // node products: c = -(Wg0c sigc + g0c Wsig) -> ggs (temporarily), d = g0c sigc -> sys scratch
  double *dtmp = A.sys + A.soff[b];
  for (int x = tid; x < r * nn; x += NT) {
    const int t = x / nn, i = (x % nn) / n, j = x % n;
    const double *f = g0c + (size_t)t * nn, *wfk = Wg0c + (size_t)t * nn, *g = sigc + (size_t)t * nn, *wg = wf + (size_t)t * nn;
    double s1 = 0.0, s2 = 0.0;
    for (int l = 0; l < n; ++l) {
      s1 += wfk[i * n + l] * g[l * n + j] + f[i * n + l] * wg[l * n + j];
      s2 += f[i * n + l] * g[l * n + j];
    }
    ggs[x] = -s1;
    dtmp[x] = s2;
  }
  __syncthreads();
  // h = beta (cf2it c + tcf2it d); ggs = -(h + eta G0) -> wf (scratch), then copy to ggs
  for (int x = tid; x < r * nn; x += NT) {
    const int t = x / nn, c = x % nn;
    double h1 = 0.0, h2 = 0.0;
    for (int k = 0; k < r; ++k) {
      h1 += A.cf2it[t * r + k] * ggs[(size_t)k * nn + c];
      h2 += A.tcf2it[t * r + k] * dtmp[(size_t)k * nn + c];
    }
    wf[x] = -(A.beta * (h1 + h2) + eta * G0[x]);
  }
  __syncthreads();
  for (int x = tid; x < r * nn; x += NT) ggs[x] = wf[x];
  __syncthreads();
  // f = vals2coefs(ggs) in place, W f -> wf
  for (int c = tid; c < nn; c += NT) lu_solve_col(A.lu, A.piv, r, ggs + c, nn);
  __syncthreads();
  for (int x = tid; x < r * nn; x += NT) {
    const int t = x / nn, c = x % nn;
    double s = 0.0;
    for (int k = 0; k < r; ++k) s += A.W[t * r + k] * ggs[(size_t)k * nn + c];
    wf[x] = s;
  }
  __syncthreads();
  // coefficient-space convolution matrix Mc
  double *S = A.sys + A.soff[b];
  for (int x = tid; x < N * N; x += NT) {
    const int row = x / N, col = x % N;
    const int t = row / n, i = row % n, kp = col / n, ip = col % n;
    const int c = i * n + ip;
    double s = -A.cf2it[t * r + kp] * wf[(size_t)kp * nn + c] + A.tcf2it[t * r + kp] * ggs[(size_t)kp * nn + c];
    for (int k = 0; k < r; ++k) s -= A.cf2it[t * r + k] * ggs[(size_t)k * nn + c] * A.W[k * r + kp];
    S[x] = A.beta * s;
  }
  __syncthreads();
  // C = Mc cf2it^{-1}: transposed solve per (row, column orbital), then the identity
  for (int x = tid; x < N * n; x += NT) lu_solve_t_col(A.lu, A.piv, r, S + (size_t)(x / n) * N + (x % n), n);
  __syncthreads();
  for (int x = tid; x < N; x += NT) S[(size_t)x * N + x] += 1.0;
}

// Phase 2 (one workgroup per block): pivoted LU of the system matrix and the solve with G0 columns.
__global__ void __launch_bounds__(256) dyson_solve_kernel(DysonBatch Bt) {
  const int pidx = find_seg(Bt.prefix, Bt.nprobs, blockIdx.x);
  const DysonArgs &A = Bt.probs[pidx];
  const int b = blockIdx.x - Bt.prefix[pidx], n = A.dims[b], r = A.r, N = r * n;
  const int tid = threadIdx.x, NT = blockDim.x;
  double *S = A.sys + A.soff[b];
  int *ip = A.ipiv + A.poff[b];
  __shared__ double red_v[256];
  __shared__ int red_i[256];
  for (int k = 0; k < N; ++k) {
    // pivot: first row of maximal |S[i][k]|, i >= k
    double best = -1.0;
    int bi = N;
    for (int i = k + tid; i < N; i += NT) {
      const double v = fabs(S[(size_t)i * N + k]);
      if (v > best || (v == best && i < bi)) { best = v; bi = i; }
    }
    red_v[tid] = best;
    red_i[tid] = bi;
    __syncthreads();
    for (int h = NT / 2; h > 0; h >>= 1) {
      if (tid < h) {
        const double v2 = red_v[tid + h];
        const int i2 = red_i[tid + h];
        if (v2 > red_v[tid] || (v2 == red_v[tid] && i2 < red_i[tid])) { red_v[tid] = v2; red_i[tid] = i2; }
      }
      __syncthreads();
    }
    const int p = red_i[0];
    if (tid == 0) ip[k] = p;
    if (p != k)
      for (int j = tid; j < N; j += NT) {
        const double a = S[(size_t)k * N + j];
        S[(size_t)k * N + j] = S[(size_t)p * N + j];
        S[(size_t)p * N + j] = a;
      }
    __syncthreads();
    const double pv = S[(size_t)k * N + k];
    for (int i = k + 1 + tid; i < N; i += NT) S[(size_t)i * N + k] /= pv;
    __syncthreads();
    const int m = N - k - 1;
    for (int x = tid; x < m * m; x += NT) {
      const int i = k + 1 + x / m, j = k + 1 + x % m;
      S[(size_t)i * N + j] -= S[(size_t)i * N + k] * S[(size_t)k * N + j];
    }
    __syncthreads();
  }
  // right-hand sides: B[(t,i)][j] = G0[t][i][j]; one thread per column j
  double *B = A.gout + A.goff[b];
  const double *G0 = A.G0 + A.goff[b];
  for (int x = tid; x < N * n; x += NT) B[x] = G0[x];
  __syncthreads();
  for (int j = tid; j < n; j += NT) {
    for (int k = 0; k < N; ++k)
      if (ip[k] != k) {
        const double a = B[(size_t)k * n + j];
        B[(size_t)k * n + j] = B[(size_t)ip[k] * n + j];
        B[(size_t)ip[k] * n + j] = a;
      }
    for (int i = 1; i < N; ++i) {
      double s = B[(size_t)i * n + j];
      for (int k = 0; k < i; ++k) s -= S[(size_t)i * N + k] * B[(size_t)k * n + j];
      B[(size_t)i * n + j] = s;
    }
    for (int i = N - 1; i >= 0; --i) {
      double s = B[(size_t)i * n + j];
      for (int k = i + 1; k < N; ++k) s -= S[(size_t)i * N + k] * B[(size_t)k * n + j];
      B[(size_t)i * n + j] = s / S[(size_t)i * N + i];
    }
  }
  __syncthreads();
  // hermitize (real symmetric part) per node
  for (int x = tid; x < r * n * n; x += NT) {
    const int t = x / (n * n), i = (x / n) % n, j = x % n;
    if (i < j) {
      const double a = B[(size_t)t * n * n + i * n + j], c = B[(size_t)t * n * n + j * n + i];
      const double h = 0.5 * (a + c);
      B[(size_t)t * n * n + i * n + j] = h;
      B[(size_t)t * n * n + j * n + i] = h;
    }
  }
}

// Per problem (one workgroup each): Z = -Tr G_new(beta) from the DLR coefficients of the diagonal
// entries (LU solve, summed in a fixed order by one thread), deta = ln|Z| / beta, eta += deta; the
// difference accumulator is cleared for the update kernel that follows.
__global__ void __launch_bounds__(256) trace_kernel(IterBatch Bt) {
  const IterArgs &A = Bt.probs[blockIdx.x];
  const int tid = threadIdx.x, r = A.r;
  for (int d = tid; d < A.ndiag; d += blockDim.x) {
    int b = 0, i = d;
    while (i >= A.dims[b]) { i -= A.dims[b]; ++b; }
    const int n = A.dims[b];
    double *x = A.trs + (size_t)d * r;
    for (int t = 0; t < r; ++t) x[t] = A.Gnew[A.goff[b] + (size_t)t * n * n + i * n + i];
    lu_solve_col(A.lu, A.piv, r, x, 1);
    double acc = 0.0;
    for (int k = 0; k < r; ++k) acc += A.kbeta[k] * x[k];
    x[0] = acc;
  }
  __syncthreads();
  if (tid == 0) {
    double tr = 0.0;
    for (int d = 0; d < A.ndiag; ++d) tr += A.trs[(size_t)d * r];
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

// Converged (diff < tol): G stays (the solver stops without mixing). Otherwise G = mix G_new + (1 - mix) G.
__global__ void __launch_bounds__(256) mix_kernel(IterBatch Bt) {
  const IterArgs &A = Bt.probs[blockIdx.y];
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool conv = A.status[kStDiff] < A.tol;
  if (x == 0) A.status[kStConv] = conv ? 1.0 : 0.0;
  if (!conv && x < A.total) A.G[x] = A.mix * A.Gnew[x] + (1.0 - A.mix) * A.G[x];
}

// max |G| per (problem, block), in a fixed order
__global__ void __launch_bounds__(256) block_max_kernel(IterBatch Bt) {
  __shared__ double part[256];
  const int pidx = find_seg(Bt.prefix, Bt.nprobs, blockIdx.x);
  const IterArgs &A = Bt.probs[pidx];
  const int b = blockIdx.x - Bt.prefix[pidx];
  double m = 0.0;
  for (int64_t x = A.goff[b] + threadIdx.x; x < A.gend[b]; x += blockDim.x) m = fmax(m, fabs(A.G[x]));
  part[threadIdx.x] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < (int)blockDim.x; ++q) t = fmax(t, part[q]);
    A.status[kStBmax + b] = t;
  }
}

__global__ void fill_kernel(int64_t n, double v, double *p) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) p[x] = v;
}

// host-callable launchers
hipError_t launch_dyson(const DysonBatch &B, hipStream_t s) {
  if (B.total <= 0) return hipSuccess;
  hipLaunchKernelGGL(dyson_build_kernel, dim3(B.total), dim3(256), 0, s, B);
  hipLaunchKernelGGL(dyson_solve_kernel, dim3(B.total), dim3(256), 0, s, B);
  return hipGetLastError();
}

hipError_t launch_iterate(const IterBatch &B, hipStream_t s) {
  if (B.nprobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(trace_kernel, dim3(B.nprobs), dim3(256), 0, s, B);
  const dim3 grid((unsigned)((B.max_total + 255) / 256), B.nprobs);
  hipLaunchKernelGGL(scale_diff_kernel, grid, dim3(256), 0, s, B);
  hipLaunchKernelGGL(mix_kernel, grid, dim3(256), 0, s, B);
  return hipGetLastError();
}

hipError_t launch_block_max(const IterBatch &B, hipStream_t s) {
  if (B.total_blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(block_max_kernel, dim3(B.total_blocks), dim3(256), 0, s, B);
  return hipGetLastError();
}

hipError_t launch_fill(int64_t n, double v, double *p, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, v, p);
  return hipGetLastError();
}

}  // namespace xca::dev

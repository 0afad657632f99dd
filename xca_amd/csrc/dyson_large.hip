// Dyson solve of blocks the one-workgroup kernel does not take (more than 12 states, or r n > 784:
// the 5-orbital shell's blocks of 25 to 100 states): the coefficient form of dyson_it_ppsc::solve
// (c++/triqs_xca/dlr_dyson_ppsc.hpp:135-168) as in dyson_kernel, restated as device-wide steps.
//   sigc = vals2coefs(Sigma_b), W sigc                       (LU solves with cf2it's factors, GEMM)
//   c = -(Wg0c sigc + g0c W sigc), d = g0c sigc              (n x n products, batched over nodes)
//   ggs = -(beta (cf2it c + tcf2it d) + eta G0)              (GEMMs along the node index)
//   f = vals2coefs(ggs), W f
//   M_(i,i') = cf2it diag(f_(i,i')) W                        (r x r products, batched over pairs)
//   S[(t,i),(k',i')] = beta (-cf2it(t,k') (W f)_k'[i,i'] + tcf2it(t,k') f_k'[i,i'] - M_(i,i')(t,k'))
//                      + delta(i,i') cf2it(t,k')
//   S g = G0 (columns j of G0_b as right-hand sides), G = (cf2it x I) g, hermitised
//   (python/triqs_xca/block_sparse_solver.py:579-587).
// The LU factorisations and the GEMMs are library calls (rocSOLVER getrf/getrs, rocBLAS GEMM and
// transposes) on the context's stream; the element-wise steps are the kernels below. Node-major
// work arrays are column-major r x n^2 (node index fastest); the packed G layout (t, i, j) is
// column-major n^2 x r. One workspace, sized for the largest such block, serves every block in turn.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/xca_amd.h"
#include "host/errors.hpp"
#include "ppsc.hpp"

namespace xca::dev {

// A[c][t + r k] = cf2it(t, k) f(k, c): the left factors of M_c = cf2it diag(f_c) W, c < n^2
__global__ void __launch_bounds__(256) pair_scale_kernel(double *__restrict__ A, const double *__restrict__ cf2it, const double *__restrict__ fcm,
                                                          int r, int64_t total) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= total) return;
  const int64_t rr = (int64_t)r * r;
  const int64_t c = x / rr;
  const int e = (int)(x % rr), t = e % r, k = e / r;
  A[x] = cf2it[t * r + k] * fcm[k + (int64_t)r * c];
}

// The N x N system, column-major (row (t, i) = t n + i, column (k', i') = k' n + i'), and the
// right-hand sides B[(t n + i) + N j] = G0_b[t][i][j].
__global__ void __launch_bounds__(256) assemble_kernel(double *__restrict__ S, double *__restrict__ B, const double *__restrict__ M,
                                                        const double *__restrict__ fcm, const double *__restrict__ wfcm,
                                                        const double *__restrict__ cf2it, const double *__restrict__ tcf2it,
                                                        const double *__restrict__ G0b, double beta, int r, int n) {
  const int N = r * n;
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t NN = (int64_t)N * N;
  if (x < NN) {
    const int row = (int)(x % N), col = (int)(x / N);
    const int t = row / n, i = row % n, kp = col / n, ip = col % n;
    const int64_t c = (int64_t)i * n + ip, rr = (int64_t)r * r;
    const double cf = cf2it[t * r + kp];
    const double sv = -cf * wfcm[kp + (int64_t)r * c] + tcf2it[t * r + kp] * fcm[kp + (int64_t)r * c] - M[c * rr + t + (int64_t)r * kp];
    S[x] = beta * sv + (i == ip ? cf : 0.0);
  } else if (x < NN + (int64_t)N * n) {
    const int64_t y = x - NN;
    const int row = (int)(y % N), j = (int)(y / N);
    const int t = row / n, i = row % n;
    B[y] = G0b[((int64_t)t * n + i) * n + j];
  }
}

// G[t][i][j] = (out_ij + out_ji) / 2 with out_ij = sum_k cf2it(t, k) g[(k n + i) + N j]
__global__ void __launch_bounds__(256) finish_kernel(double *__restrict__ G, const double *__restrict__ g, const double *__restrict__ cf2it,
                                                      int r, int n) {
  const int N = r * n;
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (int64_t)r * n * n) return;
  const int t = (int)(x / ((int64_t)n * n)), i = (int)((x / n) % n), j = (int)(x % n);
  double a = 0.0, b = 0.0;
  for (int k = 0; k < r; ++k) {
    const double cf = cf2it[t * r + k];
    a += cf * g[(int64_t)k * n + i + (int64_t)N * j];
    b += cf * g[(int64_t)k * n + j + (int64_t)N * i];
  }
  G[x] = 0.5 * (a + b);
}

}  // namespace xca::dev

namespace xca {

namespace {

void rb_check(rocblas_status st, const char *what) {
  if (st != rocblas_status_success)
    throw std::runtime_error(std::string("xca: ") + what + " failed in the large-block Dyson solve (" + rocblas_status_to_string(st) + ")");
}

unsigned grid_of(int64_t total) { return (unsigned)((total + 255) / 256); }

}  // namespace

struct LargeDyson {
  rocblas_handle h = nullptr;
  std::vector<int> blocks;                    // blocks with r n > the one-workgroup limit
  int n_max = 0;
  DevBuf<double> lu_cm, T1, T2, T3, T4, T5, T6, A, M, B;
  DevBuf<int> lu_piv, ipiv, info;
  ~LargeDyson() {
    if (h) (void)rocblas_destroy_handle(h);
  }
};

void large_dyson_release(LargeDyson *L) { delete L; }

int64_t large_dyson_bytes(int r, int n_max) {
  if (n_max <= 0) return 0;
  const int64_t N = (int64_t)r * n_max, rnn = (int64_t)r * n_max * n_max;
  return 8 * (2 * N * N + 6 * rnn + N * n_max + (int64_t)r * r) + 4 * (N + r + 2);
}

LargeDyson *large_dyson_init(xca_ctx *c, const std::vector<int> &blocks) {
  if (blocks.empty()) return nullptr;
  auto L = std::make_unique<LargeDyson>();
  L->blocks = blocks;
  const Problem &P = c->M.P;
  const int r = P.r;
  for (int b : blocks) L->n_max = std::max(L->n_max, P.dims[b]);
  const int64_t N = (int64_t)r * L->n_max, nn = (int64_t)L->n_max * L->n_max;
  if (N * N > (int64_t)1 << 31) throw std::invalid_argument("xca: Dyson system of a block too large (r * n above 46340)");
  rb_check(rocblas_create_handle(&L->h), "rocblas_create_handle");
  rb_check(rocblas_set_stream(L->h, c->stream), "rocblas_set_stream");
  // cf2it column-major, factorised on the device once (vals2coefs of every step)
  const auto &cf = c->M.ops.cf2it;
  std::vector<double> cm((size_t)r * r);
  for (int t = 0; t < r; ++t)
    for (int k = 0; k < r; ++k) cm[t + (size_t)r * k] = cf[(size_t)t * r + k];
  L->lu_cm.upload(cm, c->stream);
  L->lu_piv.alloc(r);
  L->info.alloc(2);
  rb_check(rocsolver_dgetrf(L->h, r, r, L->lu_cm.ptr, r, L->lu_piv.ptr, L->info.ptr), "rocsolver_dgetrf (cf2it)");
  for (auto *buf : {&L->T1, &L->T2, &L->T3, &L->T4, &L->T5, &L->T6}) buf->alloc((size_t)(r * nn));
  L->A.alloc((size_t)(N * N));
  L->M.alloc((size_t)(N * N));
  L->B.alloc((size_t)(N * L->n_max));
  L->ipiv.alloc((size_t)N);
  int info = 0;
  HIP_CHECK(hipMemcpyAsync(&info, L->info.ptr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  if (info != 0) throw std::runtime_error("xca: cf2it is singular (LU pivot " + std::to_string(info) + ")");
  return L.release();
}

// G_new of every large block for Sigma (device, packed) and eta.
void large_dyson_solve(xca_ctx *c, LargeDyson *L, const double *sigma, double eta, double *gnew) {
  const Problem &P = c->M.P;
  XcaDyson *d = c->dyson;
  const int r = P.r;
  const double beta = P.beta, one = 1.0, zero = 0.0, mone = -1.0, mbeta = -beta, meta = -eta;
  hipStream_t s = c->stream;
  rb_check(rocblas_set_stream(L->h, s), "rocblas_set_stream");
  const double *W = d->W.ptr, *cf = d->cf2it.ptr, *tcf = d->tcf2it.ptr;
  for (int b : L->blocks) {
    const int n = P.dims[b], nn = n * n, N = r * n;
    const int64_t go = c->M.G_off[b];
    const double *Sig = sigma + go, *G0b = d->G0.ptr + go, *g0c = d->g0c.ptr + go, *Wg0c = d->Wg0c.ptr + go;
    double *T1 = L->T1.ptr, *T2 = L->T2.ptr, *T3 = L->T3.ptr, *T4 = L->T4.ptr, *T5 = L->T5.ptr, *T6 = L->T6.ptr;
    rocblas_handle h = L->h;
    // sigc = vals2coefs(Sigma_b) and W sigc, node-major
    rb_check(rocblas_dgeam(h, rocblas_operation_transpose, rocblas_operation_none, r, nn, &one, Sig, nn, &zero, T1, r, T1, r), "transpose");
    rb_check(rocsolver_dgetrs(h, rocblas_operation_none, r, nn, L->lu_cm.ptr, r, L->lu_piv.ptr, T1, r), "rocsolver_dgetrs (sigma)");
    rb_check(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, r, nn, r, &one, W, r, T1, r, &zero, T2, r), "GEMM (W sigc)");
    // to the packed layout: per node, the n x n matrices row-major
    rb_check(rocblas_dgeam(h, rocblas_operation_transpose, rocblas_operation_none, nn, r, &one, T1, r, &zero, T3, nn, T3, nn), "transpose");
    rb_check(rocblas_dgeam(h, rocblas_operation_transpose, rocblas_operation_none, nn, r, &one, T2, r, &zero, T4, nn, T4, nn), "transpose");
    // c = -(Wg0c sigc + g0c Wsig), d = g0c sigc per node: a row-major product X Y is the
    // column-major product of the memories of Y and X
    rb_check(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, n, n, n, &mone, T3, n, nn, Wg0c, n, nn, &zero, T5, n, nn, r),
             "batched GEMM (Wg0c sigc)");
    rb_check(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, n, n, n, &mone, T4, n, nn, g0c, n, nn, &one, T5, n, nn, r),
             "batched GEMM (g0c W sigc)");
    rb_check(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, T3, n, nn, g0c, n, nn, &zero, T6, n, nn, r),
             "batched GEMM (g0c sigc)");
    // ggs = -(beta (cf2it c + tcf2it d) + eta G0), node-major
    rb_check(rocblas_dgeam(h, rocblas_operation_transpose, rocblas_operation_none, r, nn, &one, T5, nn, &zero, T1, r, T1, r), "transpose");
    rb_check(rocblas_dgeam(h, rocblas_operation_transpose, rocblas_operation_none, r, nn, &one, T6, nn, &zero, T2, r, T2, r), "transpose");
    rb_check(rocblas_dgeam(h, rocblas_operation_transpose, rocblas_operation_none, r, nn, &one, G0b, nn, &zero, T3, r, T3, r), "transpose");
    rb_check(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, r, nn, r, &mbeta, cf, r, T1, r, &zero, T4, r), "GEMM (cf2it c)");
    rb_check(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, r, nn, r, &mbeta, tcf, r, T2, r, &one, T4, r), "GEMM (tcf2it d)");
    rb_check(rocblas_daxpy(h, r * nn, &meta, T3, 1, T4, 1), "axpy (eta G0)");
    // f = vals2coefs(ggs), W f
    rb_check(rocsolver_dgetrs(h, rocblas_operation_none, r, nn, L->lu_cm.ptr, r, L->lu_piv.ptr, T4, r), "rocsolver_dgetrs (ggs)");
    rb_check(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, r, nn, r, &one, W, r, T4, r, &zero, T5, r), "GEMM (W f)");
    // M_c = (cf2it diag(f_c)) W for every orbital pair c
    const int64_t npair = (int64_t)nn * r * r;
    hipLaunchKernelGGL(dev::pair_scale_kernel, dim3(grid_of(npair)), dim3(256), 0, s, L->A.ptr, cf, T4, r, npair);
    HIP_CHECK(hipGetLastError());
    rb_check(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_transpose, r, r, r, &one, L->A.ptr, r, (rocblas_stride)r * r, W, r, 0,
                                           &zero, L->M.ptr, r, (rocblas_stride)r * r, nn),
             "batched GEMM (cf2it diag(f) W)");
    // the system (into A, free again) and the right-hand sides, then LU and the solve
    const int64_t NN = (int64_t)N * N, tot = NN + (int64_t)N * n;
    hipLaunchKernelGGL(dev::assemble_kernel, dim3(grid_of(tot)), dim3(256), 0, s, L->A.ptr, L->B.ptr, L->M.ptr, T4, T5, cf, tcf, G0b, beta, r, n);
    HIP_CHECK(hipGetLastError());
    rb_check(rocsolver_dgetrf(h, N, N, L->A.ptr, N, L->ipiv.ptr, L->info.ptr + 1), "rocsolver_dgetrf (system)");
    rb_check(rocsolver_dgetrs(h, rocblas_operation_none, N, n, L->A.ptr, N, L->ipiv.ptr, L->B.ptr, N), "rocsolver_dgetrs (system)");
    hipLaunchKernelGGL(dev::finish_kernel, dim3(grid_of((int64_t)r * nn)), dim3(256), 0, s, gnew + go, L->B.ptr, cf, r, n);
    HIP_CHECK(hipGetLastError());
    int info = 0;
    HIP_CHECK(hipMemcpyAsync(&info, L->info.ptr + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (info != 0) throw std::runtime_error("xca: the Dyson system of block " + std::to_string(b) + " is singular (LU pivot " + std::to_string(info) + ")");
  }
}

}  // namespace xca

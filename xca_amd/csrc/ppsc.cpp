// Dyson solve on the device; C ABI of the per-problem Dyson state.
// Follows dyson_it_ppsc (c++/triqs_xca/dlr_dyson_ppsc.hpp:92-168) with the hermitisation of
// BlockSparseSolver.solve_dyson (python/triqs_xca/block_sparse_solver.py:572-590).
#include "ppsc.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/xca_amd.h"
#include "host/dlr.hpp"
#include "host/errors.hpp"
#include "host/linalg.hpp"

using namespace xca;
using namespace xca::dev;

void xca_dyson_release(xca_ctx *c) {
  delete c->dyson;
  c->dyson = nullptr;
}

namespace xca {

DysonArgs dyson_args(xca_ctx *c, const double *eta_dev, const double *sigma_dev) {
  XcaDyson *d = c->dyson;
  DysonArgs A{};
  A.r = c->M.P.r; A.nb = c->M.P.nb; A.beta = c->M.P.beta; A.eta = eta_dev;
  A.dims = d->dims.ptr; A.goff = d->goff.ptr; A.soff = d->soff.ptr; A.poff = d->poff.ptr; A.roff = d->roff.ptr; A.lu = d->lu.ptr; A.piv = d->piv.ptr;
  A.cf2it = d->cf2it.ptr; A.tcf2it = d->tcf2it.ptr; A.W = d->W.ptr;
  A.G0 = d->G0.ptr; A.g0c = d->g0c.ptr; A.Wg0c = d->Wg0c.ptr; A.sigma = sigma_dev;
  A.ggs = d->ggs.ptr; A.wf = d->wf.ptr; A.sys = d->sys.ptr; A.rhs = d->rhs.ptr; A.gout = d->gnew.ptr; A.ipiv = d->ipiv.ptr;
  A.large = d->large ? d->large_flag.ptr : nullptr;
  return A;
}

IterArgs iter_args(xca_ctx *c, double *status, double tol, double mix) {
  XcaDyson *d = c->dyson;
  IterArgs A{};
  A.r = c->M.P.r; A.nb = c->M.P.nb; A.ndiag = d->ndiag;
  A.beta = c->M.P.beta; A.tol = tol; A.mix = mix; A.total = c->M.G_size;
  A.dims = d->dims.ptr; A.goff = d->goff.ptr; A.gend = d->gend.ptr; A.lu = d->lu.ptr; A.piv = d->piv.ptr;
  A.zw = d->zw.ptr; A.tau = d->tau.ptr; A.G = c->G.ptr; A.Gnew = d->gnew.ptr; A.status = status;
  return A;
}

int dyson_np_max(const xca_ctx *c) {
  int np = 16;
  for (int b = 0; b < c->M.P.nb; ++b) np = std::max(np, (c->M.P.r * c->M.P.dims[b] + 15) / 16 * 16);
  return np;
}

// Blocks the one-workgroup kernel cannot take: a padded system above kDysonWgMax rows, or more than
// 12 rows per block (its LDS panel of padded(r n) x 16 doubles holds the r n^2 vals2coefs columns
// only while n <= 16; 12 is the tile engine's block limit).
static bool is_large(const Problem &P, int b) {
  return P.dims[b] > 12 || ((int64_t)P.r * P.dims[b] + 15) / 16 * 16 > kDysonWgMax;
}

int dyson_np_max_small(const xca_ctx *c) {
  int np = 16;
  for (int b = 0; b < c->M.P.nb; ++b)
    if (!is_large(c->M.P, b)) np = std::max(np, (c->M.P.r * c->M.P.dims[b] + 15) / 16 * 16);
  return np;
}

int64_t dyson_bytes(const xca_ctx *c) {
  const Problem &P = c->M.P;
  int64_t sys = 0, rhs = 0, piv = 0, ndiag = 0;
  int n_large = 0;
  for (int b = 0; b < P.nb; ++b) {
    ndiag += P.dims[b];
    if (is_large(P, b)) {
      n_large = std::max(n_large, P.dims[b]);
      continue;
    }
    const int64_t Np = ((int64_t)P.r * P.dims[b] + 15) / 16 * 16;
    sys += Np * Np;
    rhs += Np * P.dims[b];
    piv += Np;
  }
  return 8 * (6 * c->M.G_size + sys + rhs + 6 * (int64_t)P.r * P.r + 2 * P.r) + 4 * piv + large_dyson_bytes(P.r, n_large);
}

void require_dyson(xca_ctx *c) {
  if (!c->dyson) throw std::runtime_error("xca: xca_dyson_init has not been called");
}

}  // namespace xca

extern "C" int xca_dyson_init(xca_ctx *c, const double *G0) {
  try {
    HIP_CHECK(hipSetDevice(c->device));
    xca_dyson_release(c);
    auto *d = new XcaDyson();
    c->dyson = d;
    const Problem &P = c->M.P;
    const auto &o = c->M.ops;
    const int r = P.r, nb = P.nb;
    hipStream_t s = c->stream;
    std::vector<int> dims(P.dims);
    std::vector<int64_t> goff(nb), gend(nb), soff(nb), poff(nb), roff(nb);
    std::vector<int> large_blocks, large_flag(nb, 0);
    int64_t so = 0, po = 0, ro = 0;
    for (int b = 0; b < nb; ++b) {
      // blocks whose system exceeds one workgroup are solved device-wide and hold no slot here
      const bool big = is_large(P, b);
      const int64_t Np = big ? 0 : ((int64_t)r * P.dims[b] + 15) / 16 * 16;
      if (big) {
        large_blocks.push_back(b);
        large_flag[b] = 1;
      }
      goff[b] = c->M.G_off[b];
      gend[b] = goff[b] + (int64_t)r * P.dims[b] * P.dims[b];
      soff[b] = so;
      so += Np * Np;
      poff[b] = po;
      po += Np;
      roff[b] = ro;
      ro += Np * P.dims[b];
      d->ndiag += P.dims[b];
    }
    d->dims.upload(dims, s);
    d->goff.upload(goff, s);
    d->gend.upload(gend, s);
    d->soff.upload(soff, s);
    d->poff.upload(poff, s);
    d->roff.upload(roff, s);
    d->lu.upload(o.lu, s);
    d->piv.upload(o.piv, s);
    d->cf2it.upload(o.cf2it, s);
    d->tcf2it.upload(o.tcf2it, s);
    d->W.upload(o.W, s);
    std::vector<double> tau(r), zw(r);
    for (int t = 0; t < r; ++t) tau[t] = dlr::rel2abs(P.it[t]) * P.beta;
    // Tr G(beta) = K(1, w)^T cf2it^{-1} Tr G(t) = zw . Tr G(t) with cf2it^T zw = K(1, w), solved through
    // the LU factors (P cf2it = L U): U^T y = K(1, w), L^T z = y, zw = P^T z. cf2it is far from
    // orthogonal (condition ~1e10 at eps = 1e-10): a transposed solve keeps zw . g as accurate as
    // the evaluation of vals2coefs(g) at beta, where an explicit inverse would not.
    for (int k = 0; k < r; ++k) zw[k] = dlr::kernel(1.0, P.rf[k]);
    for (int i = 0; i < r; ++i) {
      for (int k = 0; k < i; ++k) zw[i] -= o.lu[(size_t)k * r + i] * zw[k];
      zw[i] /= o.lu[(size_t)i * r + i];
    }
    for (int i = r - 1; i >= 0; --i)
      for (int k = i + 1; k < r; ++k) zw[i] -= o.lu[(size_t)k * r + i] * zw[k];
    for (int k = r - 1; k >= 0; --k)
      if (o.piv[k] != k) std::swap(zw[k], zw[o.piv[k]]);
    d->tau.upload(tau, s);
    d->zw.upload(zw, s);
    // G0 and its coefficients (dyson_it_ppsc ctor: g0c = vals2coefs(g0))
    d->G0_h.assign(G0, G0 + c->M.G_size);
    std::vector<double> g0c(d->G0_h), Wg0c(c->M.G_size);
    for (int b = 0; b < nb; ++b) {
      int n = P.dims[b];
      o.vals2coefs(g0c.data() + goff[b], n * n);
      la::matmul(o.W.data(), g0c.data() + goff[b], Wg0c.data() + goff[b], r, r, n * n);
    }
    d->G0.upload(d->G0_h, s);
    d->g0c.upload(g0c, s);
    d->Wg0c.upload(Wg0c, s);
    d->ggs.alloc(c->M.G_size);
    d->wf.alloc(c->M.G_size);
    d->gnew.alloc(c->M.G_size);
    d->sys.alloc(so);
    d->rhs.alloc(ro);
    d->ipiv.alloc(po);
    d->status.alloc(status_stride(nb));
    d->large_flag.upload(large_flag, s);
    d->large = large_dyson_init(c, large_blocks);
    d->device_bytes = dyson_bytes(c);
    HIP_CHECK(hipStreamSynchronize(s));
    return 0;
  } catch (std::exception &e) {
    xca::set_error(e.what());
    return -1;
  }
}

// G = Dyson(Sigma, eta) for every block (hermitized). sigma == NULL uses the device Sigma accumulator.
extern "C" int xca_dyson_solve(xca_ctx *c, const double *sigma, double eta, double *G_out) {
  try {
    require_dyson(c);
    HIP_CHECK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    XcaDyson *d = c->dyson;
    const double *sig = c->sigma.ptr;
    if (sigma) {
      d->sig_in.alloc(c->M.G_size);
      HIP_CHECK(hipMemcpyAsync(d->sig_in.ptr, sigma, c->M.G_size * sizeof(double), hipMemcpyHostToDevice, s));
      sig = d->sig_in.ptr;
    }
    double *eta_dev = d->status.ptr + kStEta;
    HIP_CHECK(launch_fill(1, eta, eta_dev, s));
    std::vector<DysonArgs> A(1, dyson_args(c, eta_dev, sig));
    std::vector<int> pre{0, c->M.P.nb};
    DevBuf<DysonArgs> dA;
    DevBuf<int> dpre;
    dA.upload(A, s);
    dpre.upload(pre, s);
    HIP_CHECK(launch_dyson(DysonBatch{dA.ptr, dpre.ptr, 1, c->M.P.nb, dyson_np_max_small(c)}, s));
    if (d->large) large_dyson_solve(c, d->large, sig, eta, d->gnew.ptr);
    HIP_CHECK(hipMemcpyAsync(G_out, d->gnew.ptr, c->M.G_size * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return 0;
  } catch (std::exception &e) {
    xca::set_error(e.what());
    return -1;
  }
}

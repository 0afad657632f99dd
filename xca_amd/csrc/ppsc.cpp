// Dyson solve and pseudo-particle self-consistency loop on the device; C ABI.
// Follows dyson_it_ppsc (c++/triqs_xca/dlr_dyson_ppsc.hpp:92-168) and BlockSparseSolver.solve with
// classic normalization (python/triqs_xca/block_sparse_solver.py:295-346, 427-449, 572-590).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/xca_amd.h"
#include "ctx.hpp"
#include "host/dlr.hpp"
#include "host/errors.hpp"
#include "host/linalg.hpp"

namespace xca::dev {
struct DysonArgs {
  int r, nb;
  double beta, eta;
  const int *dims;
  const int64_t *goff;
  const int64_t *soff;
  const double *lu;
  const int *piv;
  const double *cf2it, *tcf2it, *W;
  const double *G0, *g0c, *Wg0c;
  const double *sigma;
  double *sigc, *ggs, *wf, *sys, *gout;
  int *ipiv;
};
hipError_t launch_dyson(const DysonArgs &A, hipStream_t s);
hipError_t launch_trace_beta(int r, int nb, int ndiag, const int *dims, const int64_t *goff, const double *lu, const int *piv, const double *kbeta,
                             const double *G, double *scratch, double *out, hipStream_t s);
hipError_t launch_scale_diff(int64_t total, int r, int nb, const int *dims, const int64_t *goff, const double *tau, double deta, double *Gnew,
                             const double *G, unsigned long long *dmax, hipStream_t s);
hipError_t launch_mix(int64_t total, double mix, const double *Gnew, double *G, hipStream_t s);
hipError_t launch_block_max(int nb, const int64_t *goff, const int64_t *gend, const double *G, double *out, hipStream_t s);
hipError_t launch_fill(int64_t n, double v, double *p, hipStream_t s);
}  // namespace xca::dev

using namespace xca;
using namespace xca::dev;

struct XcaDyson {
  DevBuf<int> dims, piv, ipiv;
  DevBuf<int64_t> goff, gend, soff;
  DevBuf<double> lu, cf2it, tcf2it, W, tau, kbeta;
  DevBuf<double> G0, g0c, Wg0c, sigc, ggs, wf, sys, gnew, trs, scal, sig_in;
  DevBuf<unsigned long long> dmax;
  std::vector<double> G0_h;
  int ndiag = 0;
};

void xca_dyson_release(xca_ctx *c) {
  delete c->dyson;
  c->dyson = nullptr;
}

namespace {

DysonArgs dyson_args(xca_ctx *c, double eta, const double *sigma_dev) {
  XcaDyson *d = c->dyson;
  DysonArgs A{};
  A.r = c->M.P.r; A.nb = c->M.P.nb; A.beta = c->M.P.beta; A.eta = eta;
  A.dims = d->dims.ptr; A.goff = d->goff.ptr; A.soff = d->soff.ptr; A.lu = d->lu.ptr; A.piv = d->piv.ptr;
  A.cf2it = d->cf2it.ptr; A.tcf2it = d->tcf2it.ptr; A.W = d->W.ptr;
  A.G0 = d->G0.ptr; A.g0c = d->g0c.ptr; A.Wg0c = d->Wg0c.ptr; A.sigma = sigma_dev;
  A.sigc = d->sigc.ptr; A.ggs = d->ggs.ptr; A.wf = d->wf.ptr; A.sys = d->sys.ptr; A.gout = d->gnew.ptr; A.ipiv = d->ipiv.ptr;
  return A;
}

void require_dyson(xca_ctx *c) {
  if (!c->dyson) throw std::runtime_error("xca: xca_dyson_init has not been called");
}

}  // namespace

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
    std::vector<int64_t> goff(nb), gend(nb), soff(nb);
    int64_t so = 0;
    for (int b = 0; b < nb; ++b) {
      goff[b] = c->M.G_off[b];
      gend[b] = goff[b] + (int64_t)r * P.dims[b] * P.dims[b];
      soff[b] = so;
      so += (int64_t)r * P.dims[b] * r * P.dims[b];
      d->ndiag += P.dims[b];
    }
    d->dims.upload(dims, s);
    d->goff.upload(goff, s);
    d->gend.upload(gend, s);
    d->soff.upload(soff, s);
    d->lu.upload(o.lu, s);
    d->piv.upload(o.piv, s);
    d->cf2it.upload(o.cf2it, s);
    d->tcf2it.upload(o.tcf2it, s);
    d->W.upload(o.W, s);
    std::vector<double> tau(r), kbeta(r);
    for (int t = 0; t < r; ++t) tau[t] = dlr::rel2abs(P.it[t]) * P.beta;
    for (int k = 0; k < r; ++k) kbeta[k] = dlr::kernel(1.0, P.rf[k]);
    d->tau.upload(tau, s);
    d->kbeta.upload(kbeta, s);
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
    d->sigc.alloc(c->M.G_size);
    d->ggs.alloc(c->M.G_size);
    d->wf.alloc(c->M.G_size);
    d->gnew.alloc(c->M.G_size);
    d->sys.alloc(so);
    d->ipiv.alloc((size_t)r * d->ndiag);
    d->trs.alloc((size_t)r * d->ndiag);
    d->scal.alloc(std::max(8, nb));
    d->dmax.alloc(1);
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
    const double *sig = c->sigma.ptr;
    if (sigma) {
      c->dyson->sig_in.alloc(c->M.G_size);
      HIP_CHECK(hipMemcpyAsync(c->dyson->sig_in.ptr, sigma, c->M.G_size * sizeof(double), hipMemcpyHostToDevice, s));
      sig = c->dyson->sig_in.ptr;
    }
    DysonArgs A = dyson_args(c, eta, sig);
    HIP_CHECK(launch_dyson(A, s));
    HIP_CHECK(hipMemcpyAsync(G_out, c->dyson->gnew.ptr, c->M.G_size * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return 0;
  } catch (std::exception &e) {
    xca::set_error(e.what());
    return -1;
  }
}

namespace xca {
void sigma_full_order(xca_ctx *c, int max_order);  // context.cpp
}

// Pseudo-particle self-consistency: Sigma (orders 1..max_order) -> Dyson -> normalize -> mix, until
// max |G_new - G| < tol or maxiter iterations. On return G_out holds the solver's G (the last mixed
// iterate, block_sparse_solver.py:338-346), eta_out the pseudo-particle chemical potential.
extern "C" int xca_ppsc_solve(xca_ctx *c, double eta0, int max_order, double tol, int maxiter, double mix, double *G_out, double *eta_out,
                              int *iters_out, double *diff_out) {
  try {
    require_dyson(c);
    HIP_CHECK(hipSetDevice(c->device));
    XcaDyson *d = c->dyson;
    const Problem &P = c->M.P;
    hipStream_t s = c->stream;
    const int64_t total = c->M.G_size;
    HIP_CHECK(hipMemcpyAsync(c->G.ptr, d->G0.ptr, total * sizeof(double), hipMemcpyDeviceToDevice, s));
    double eta = eta0, diff = 0.0;
    int iter = 0;
    std::vector<double> bmax(P.nb);
    for (iter = 1; iter <= maxiter; ++iter) {
      // zero-block flags of G (block_sparse.cpp:45-49)
      HIP_CHECK(launch_block_max(P.nb, d->goff.ptr, d->gend.ptr, c->G.ptr, d->scal.ptr, s));
      HIP_CHECK(hipMemcpyAsync(bmax.data(), d->scal.ptr, P.nb * sizeof(double), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      c->G_zero.assign(P.nb, 0);
      for (int b = 0; b < P.nb; ++b) c->G_zero[b] = !(bmax[b] >= 1e-16);
      c->have_G = true;
      sigma_full_order(c, max_order);
      DysonArgs A = dyson_args(c, eta, c->sigma.ptr);
      HIP_CHECK(launch_dyson(A, s));
      double trz = 0.0;
      HIP_CHECK(launch_trace_beta(P.r, P.nb, d->ndiag, d->dims.ptr, d->goff.ptr, d->lu.ptr, d->piv.ptr, d->kbeta.ptr, d->gnew.ptr, d->trs.ptr,
                                  d->scal.ptr, s));
      HIP_CHECK(hipMemcpyAsync(&trz, d->scal.ptr, sizeof(double), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      const double Z = -trz;
      const double deta = std::log(std::fabs(Z)) / P.beta;
      HIP_CHECK(launch_fill(1, 0.0, reinterpret_cast<double *>(d->dmax.ptr), s));
      HIP_CHECK(launch_scale_diff(total, P.r, P.nb, d->dims.ptr, d->goff.ptr, d->tau.ptr, deta, d->gnew.ptr, c->G.ptr, d->dmax.ptr, s));
      unsigned long long dm = 0;
      HIP_CHECK(hipMemcpyAsync(&dm, d->dmax.ptr, sizeof(dm), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      std::memcpy(&diff, &dm, sizeof(double));
      eta += deta;
      if (diff < tol) break;
      HIP_CHECK(launch_mix(total, mix, d->gnew.ptr, c->G.ptr, s));
    }
    if (iter > maxiter) iter = maxiter;
    HIP_CHECK(hipMemcpyAsync(G_out, c->G.ptr, total * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (eta_out) *eta_out = eta;
    if (iters_out) *iters_out = iter;
    if (diff_out) *diff_out = diff;
    return 0;
  } catch (std::exception &e) {
    xca::set_error(e.what());
    return -1;
  }
}

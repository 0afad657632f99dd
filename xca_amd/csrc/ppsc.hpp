// Per-problem Dyson state and the descriptors the batched iteration engine launches with.
#pragma once
#include <vector>

#include "ctx.hpp"
#include "dyson.hpp"

namespace xca {
struct LargeDyson;  // blocks whose system exceeds one workgroup (dyson_large.hip)
void large_dyson_release(LargeDyson *L);
}  // namespace xca

struct XcaDyson {
  ~XcaDyson() { xca::large_dyson_release(large); }
  xca::LargeDyson *large = nullptr;
  xca::DevBuf<int> large_flag;  // (nb) 1 for the blocks of `large`
  xca::DevBuf<int> dims, piv, ipiv;
  xca::DevBuf<int64_t> goff, gend, soff, poff, roff;
  xca::DevBuf<double> lu, cf2it, tcf2it, W, tau, zw;
  xca::DevBuf<double> G0, g0c, Wg0c, ggs, wf, sys, rhs, gnew, sig_in;
  xca::DevBuf<double> status;  // iteration status of the single-problem loop and the eta scalar of xca_dyson_solve
  std::vector<double> G0_h;
  int ndiag = 0;
  int64_t device_bytes = 0;
};

namespace xca {
// Dyson descriptor of one problem; eta is a device scalar, sigma a device buffer.
dev::DysonArgs dyson_args(xca_ctx *c, const double *eta_dev, const double *sigma_dev);
// Normalisation / convergence / mixing descriptor of one problem writing its status at `status`.
dev::IterArgs iter_args(xca_ctx *c, double *status, double tol, double mix);
// Largest padded Dyson system dimension (r n rounded up to 16) over the blocks.
int dyson_np_max(const xca_ctx *c);
// The same over the blocks the one-workgroup kernel solves (padded r n at most kDysonWgMax).
int dyson_np_max_small(const xca_ctx *c);
constexpr int kDysonWgMax = 784;
LargeDyson *large_dyson_init(xca_ctx *c, const std::vector<int> &blocks);
void large_dyson_solve(xca_ctx *c, LargeDyson *L, const double *sigma, double eta, double *gnew);
int64_t large_dyson_bytes(int r, int n_max);
// Device bytes xca_dyson_init allocates for this problem (without allocating).
int64_t dyson_bytes(const xca_ctx *c);
void require_dyson(xca_ctx *c);
}  // namespace xca

// Evaluator context shared by the host orchestration files.
#pragma once
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine.hpp"
#include "host/model.hpp"

namespace xca {

#define HIP_CHECK(x)                                                                                          \
  do {                                                                                                        \
    hipError_t e_ = (x);                                                                                      \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
  } while (0)

template <class T> struct DevBuf {
  T *ptr = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count <= n && ptr) return;
    release();
    n = std::max<size_t>(count, 1);
    HIP_CHECK(hipMalloc(&ptr, n * sizeof(T)));
  }
  void upload(const std::vector<T> &v, hipStream_t s) {
    alloc(v.size());
    if (!v.empty()) HIP_CHECK(hipMemcpyAsync(ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  }
};

struct DevPlan;
struct CorrDevPlan;
struct WidePlan;
struct WideCtx;
struct SigmaAll;

}  // namespace xca

// Dyson / iteration state (ppsc.cpp)
struct XcaDyson;

struct xca_ctx {
  using Plans = std::map<std::vector<int>, std::unique_ptr<xca::DevPlan>>;
  xca::Model M;
  xca::PoleSets ps;
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int RB = 0, RP = 0, maxexp = 8;
  int tile_variant = -1;  // forced tile variant, -1: autotuned per plan
  int launch_mode = 0;    // 0: one persistent chain launch per plan; 1: one chain launch per chunk
  int last_variant = -1;
  std::vector<int> plist_h;
  int plist_off[4];  // fwd_a, fwd_b, bwd_a, bwd_b
  xca::DevBuf<double> opmats, kp, km, delta, delta_refl, kpow, G, sigma, frag;
  xca::DevBuf<int> plist;
  xca::DevBuf<int64_t> sig_off;
  xca::DevBuf<int> sig_size;
  xca::dev::DlrFrag D{};
  std::vector<int> G_zero;
  bool have_G = false;
  Plans plans;
  // zero-block patterns that have plans, least recently used first; at most kMaxZeroPatterns are
  // kept, so a G whose zero blocks keep changing does not accumulate plans
  std::vector<std::vector<int>> zero_lru;
  static constexpr int kMaxZeroPatterns = 4;
  std::map<std::vector<int>, std::unique_ptr<xca::CorrDevPlan>> corr_plans;
  bool wide = false;  // blocks above the tile engine's 12 rows: the large-block engine evaluates Sigma
  bool outer_events = false;  // ev0/ev1 bracket a sequence of evaluations: inner launches leave them
  std::map<std::vector<int>, std::unique_ptr<xca::WidePlan>> wide_plans;
  std::unique_ptr<xca::WideCtx> widectx;
  std::unique_ptr<xca::SigmaAll> sigall;
  xca::DevBuf<double> corr_out;
  XcaDyson *dyson = nullptr;
  int64_t last_items = 0;
  double last_macs = 0, last_ms = 0;
  std::vector<double> ppsc_history;  // (diff, Z, eta) per iteration of the last xca_ppsc_solve
};

void xca_dyson_release(xca_ctx *c);

// Dyson solve that allocates nothing: xca_dyson_solve_noalloc. Everything the Dyson kernels read or
// write beyond the caller's arrays (descriptors, DLR factors, G0 and its coefficients, the node
// products, the block systems, right-hand sides, pivots and the output staging) lives in one
// file-scope device array of fixed size, so the call makes no device allocation.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/xca_amd.h"
#include "ctx.hpp"
#include "dyson.hpp"
#include "host/errors.hpp"
#include "host/linalg.hpp"
#include "ppsc.hpp"

namespace {

constexpr size_t kScratchBytes = size_t(32) << 20;  // 32 MiB per GPU, held until the process ends
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

// Device address of each GPU's copy of `scratch`, looked up once per device with that device current.
std::mutex g_lock;
void *g_ptr[kMaxDevices] = {};

char *scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) throw std::runtime_error("xca_dyson_solve_noalloc: device index out of range");
  std::lock_guard<std::mutex> guard(g_lock);
  if (!g_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) throw std::runtime_error(std::string("xca_dyson_solve_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    g_ptr[device] = p;
  }
  return static_cast<char *>(g_ptr[device]);
}

// Bump layout of the scratch array: every piece starts on a 256-byte boundary.
struct Layout {
  size_t used = 0;
  size_t take(size_t bytes) {
    const size_t off = used;
    used += (bytes + 255) / 256 * 256;
    return off;
  }
};

}  // namespace

using namespace xca;
using namespace xca::dev;

// G = Dyson(Sigma, eta) for every block, hermitised, as xca_dyson_solve, with G0 given here instead
// of by xca_dyson_init.
extern "C" int xca_dyson_solve_noalloc(xca_ctx *c, const double *G0, const double *sigma, double eta, double *G_out) {
  try {
    if (!c) throw std::invalid_argument("xca_dyson_solve_noalloc: null context");
    if (!G0 || !sigma || !G_out) throw std::invalid_argument("xca_dyson_solve_noalloc: G0, sigma and G_out are required");
    HIP_CHECK(hipSetDevice(c->device));
    const Problem &P = c->M.P;
    const auto &o = c->M.ops;
    const int r = P.r, nb = P.nb;
    const int64_t gs = c->M.G_size;
    const int np_max = dyson_np_max(c);
    if (np_max > 784) throw std::invalid_argument("xca: the Dyson solve supports r * (block dimension) <= 784");
    std::vector<int64_t> goff(nb), soff(nb), poff(nb), roff(nb);
    int64_t so = 0, po = 0, ro = 0;
    for (int b = 0; b < nb; ++b) {
      const int64_t Np = ((int64_t)r * P.dims[b] + 15) / 16 * 16;
      goff[b] = c->M.G_off[b];
      soff[b] = so;
      so += Np * Np;
      poff[b] = po;
      po += Np;
      roff[b] = ro;
      ro += Np * P.dims[b];
    }
    const bool sys_lds = np_max <= kDysonLdsSysMax;  // the kernel then keeps every system in LDS
    Layout L;
    const size_t oA = L.take(sizeof(DysonArgs)), oPre = L.take(2 * sizeof(int)), oDims = L.take(nb * sizeof(int));
    const size_t oGoff = L.take(nb * sizeof(int64_t)), oSoff = L.take(nb * sizeof(int64_t)), oPoff = L.take(nb * sizeof(int64_t)),
                 oRoff = L.take(nb * sizeof(int64_t));
    const size_t oLu = L.take((size_t)r * r * 8), oPiv = L.take(r * sizeof(int)), oCf = L.take((size_t)r * r * 8),
                 oTcf = L.take((size_t)r * r * 8), oW = L.take((size_t)r * r * 8);
    const size_t oG0 = L.take(gs * 8), oG0c = L.take(gs * 8), oWg0c = L.take(gs * 8), oSig = L.take(gs * 8), oGgs = L.take(gs * 8),
                 oWf = L.take(gs * 8), oOut = L.take(gs * 8);
    const size_t oSys = L.take(sys_lds ? 8 : (size_t)so * 8), oRhs = L.take((size_t)ro * 8), oIp = L.take((size_t)po * sizeof(int));
    const size_t oEta = L.take(8);
    if (L.used > kScratchBytes)
      throw std::invalid_argument("xca_dyson_solve_noalloc: the problem needs " + std::to_string(L.used) +
                                  " bytes of device scratch, above the static maximum of " + std::to_string(kScratchBytes) +
                                  " bytes (use xca_dyson_init / xca_dyson_solve)");
    char *base = scratch_on(c->device);
    auto dptr = [&](size_t off) { return reinterpret_cast<double *>(base + off); };
    // G0's DLR coefficients and W applied to them (dyson_it_ppsc ctor: g0c = vals2coefs(g0))
    std::vector<double> g0c(G0, G0 + gs), Wg0c(gs);
    for (int b = 0; b < nb; ++b) {
      const int n = P.dims[b];
      o.vals2coefs(g0c.data() + goff[b], n * n);
      la::matmul(o.W.data(), g0c.data() + goff[b], Wg0c.data() + goff[b], r, r, n * n);
    }
    DysonArgs A{};
    A.r = r; A.nb = nb; A.beta = P.beta; A.eta = dptr(oEta);
    A.dims = reinterpret_cast<const int *>(base + oDims);
    A.goff = reinterpret_cast<const int64_t *>(base + oGoff);
    A.soff = reinterpret_cast<const int64_t *>(base + oSoff);
    A.poff = reinterpret_cast<const int64_t *>(base + oPoff);
    A.roff = reinterpret_cast<const int64_t *>(base + oRoff);
    A.lu = dptr(oLu); A.piv = reinterpret_cast<const int *>(base + oPiv);
    A.cf2it = dptr(oCf); A.tcf2it = dptr(oTcf); A.W = dptr(oW);
    A.G0 = dptr(oG0); A.g0c = dptr(oG0c); A.Wg0c = dptr(oWg0c); A.sigma = dptr(oSig);
    A.ggs = dptr(oGgs); A.wf = dptr(oWf); A.sys = dptr(oSys); A.rhs = dptr(oRhs); A.gout = dptr(oOut);
    A.ipiv = reinterpret_cast<int *>(base + oIp);
    const int pre[2] = {0, nb};
    hipStream_t s = c->stream;
    auto put = [&](size_t off, const void *src, size_t bytes) {
      if (bytes) HIP_CHECK(hipMemcpyAsync(base + off, src, bytes, hipMemcpyHostToDevice, s));
    };
    put(oA, &A, sizeof(A));
    put(oPre, pre, sizeof(pre));
    put(oDims, P.dims.data(), nb * sizeof(int));
    put(oGoff, goff.data(), nb * sizeof(int64_t));
    put(oSoff, soff.data(), nb * sizeof(int64_t));
    put(oPoff, poff.data(), nb * sizeof(int64_t));
    put(oRoff, roff.data(), nb * sizeof(int64_t));
    put(oLu, o.lu.data(), (size_t)r * r * 8);
    put(oPiv, o.piv.data(), r * sizeof(int));
    put(oCf, o.cf2it.data(), (size_t)r * r * 8);
    put(oTcf, o.tcf2it.data(), (size_t)r * r * 8);
    put(oW, o.W.data(), (size_t)r * r * 8);
    put(oG0, G0, gs * 8);
    put(oG0c, g0c.data(), gs * 8);
    put(oWg0c, Wg0c.data(), gs * 8);
    put(oSig, sigma, gs * 8);
    put(oEta, &eta, 8);
    HIP_CHECK(launch_dyson(DysonBatch{reinterpret_cast<const DysonArgs *>(base + oA), reinterpret_cast<const int *>(base + oPre), 1, nb, np_max}, s));
    HIP_CHECK(hipMemcpyAsync(G_out, base + oOut, gs * 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return 0;
  } catch (std::exception &e) {
    xca::set_error(e.what());
    return -1;
  }
}

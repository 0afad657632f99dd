// C ABI: DLR grid construction and kernel evaluation (host only).
#include <cstring>
#include <exception>
#include <string>

#include "../../include/xca_amd.h"
#include "host/dlr.hpp"
#include "host/errors.hpp"

extern "C" int xca_dlr_build(double lambda, double eps, int max_r, double *rf_out, double *it_out) {
  try {
    auto rf = xca::dlr::build_rf(lambda, eps);
    auto it = xca::dlr::build_it(lambda, rf);
    int r = (int)rf.size();
    if (r > max_r) {
      xca::set_error("xca_dlr_build: rank " + std::to_string(r) + " exceeds max_r = " + std::to_string(max_r));
      return -1;
    }
    std::memcpy(rf_out, rf.data(), sizeof(double) * r);
    std::memcpy(it_out, it.data(), sizeof(double) * r);
    return r;
  } catch (std::exception &e) {
    xca::set_error(e.what());
    return -1;
  }
}

extern "C" double xca_dlr_kernel(double t, double om) { return xca::dlr::kernel(t, om); }

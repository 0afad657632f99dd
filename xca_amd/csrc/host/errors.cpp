#include "errors.hpp"

#include "../../../include/xca_amd.h"

namespace xca {
static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }
const char *get_error() { return g_error.c_str(); }
} // namespace xca

extern "C" const char *xca_last_error(void) { return xca::get_error(); }

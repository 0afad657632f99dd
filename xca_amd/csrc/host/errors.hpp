// Per-thread error message behind xca_last_error().
#pragma once
#include <string>

namespace xca {
void set_error(const std::string &msg);
const char *get_error();
} // namespace xca

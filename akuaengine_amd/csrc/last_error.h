// last_error.h -- the per-thread message behind akua_last_error(), shared by the library's
// source files (the storage lives in solver.cpp).
#pragma once

#include <string>

namespace akua {

void set_last_error(const std::string& message);

}  // namespace akua

// Shared host-side helpers of the cbx library (error state, checks).
#pragma once

#include "../../include/cbx.h"

#include <cstdio>
#include <string>

namespace cbx {

// Records `msg` as the calling thread's last error and returns `code`.
cbx_status set_error(cbx_status code, const char* fmt, ...);

}  // namespace cbx

#define CBX_RETURN_IF(cond, code, ...)                \
  do {                                                \
    if (cond) return ::cbx::set_error(code, __VA_ARGS__); \
  } while (0)

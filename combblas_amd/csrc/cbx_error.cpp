// Thread-local error text behind cbx_last_error().
#include "cbx_internal.h"

#include <cstdarg>
#include <cstdlib>

namespace cbx {

static thread_local char g_last_error[512] = "";

cbx_status set_error(cbx_status code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

}  // namespace cbx

extern "C" const char* cbx_last_error(void) { return cbx::g_last_error; }

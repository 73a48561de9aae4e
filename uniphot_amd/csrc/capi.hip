// C-ABI utilities: version and the thread-local last-error string.
#include <stdarg.h>
#include <stdio.h>

#include "uph_capi_internal.hpp"

namespace uph {
static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
}  // namespace uph

extern "C" int uph_version(void) { return 100; }

extern "C" const char* uph_last_error(void) { return uph::g_err; }

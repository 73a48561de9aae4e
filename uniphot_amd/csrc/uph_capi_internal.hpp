// Status codes and the thread-local last-error string of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdio.h>

#include "../../include/uniphot_amd.h"

namespace uph {

void set_error(const char* fmt, ...);

inline int hip_status(hipError_t e, const char* what) {
  if (e == hipSuccess) return UPH_OK;
  set_error("%s: %s", what, hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? UPH_ERR_OOM : UPH_ERR_INTERNAL;
}

#define UPH_CHECK_ARG(cond, ...)        \
  do {                                  \
    if (!(cond)) {                      \
      ::uph::set_error(__VA_ARGS__);    \
      return UPH_ERR_INVALID_ARGUMENT;  \
    }                                   \
  } while (0)

}  // namespace uph

// Shared definitions for the yastn_hip C-ABI library (gfx950 / CDNA4 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/yastn_hip.h"

namespace yh {

// Set the thread-local error message returned by yastn_hip_last_error().
void set_error(const char* fmt, ...);

// Record a HIP runtime error (if any) and map it to a status code.
int hip_status(hipError_t e, const char* where);

inline int itemsize_of(int dtype) { return dtype == YH_C128 ? 16 : (dtype == YH_F64 ? 8 : 0); }

__host__ __device__ inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace yh

#define YH_CHECK_ARG(cond, ...)          \
  do {                                   \
    if (!(cond)) {                       \
      yh::set_error(__VA_ARGS__);        \
      return YH_ERR_ARG;                 \
    }                                    \
  } while (0)

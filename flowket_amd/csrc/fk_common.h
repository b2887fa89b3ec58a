// Shared helpers for the flowket_amd HIP library (gfx950 / CDNA4 only).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/flowket_hip.h"

namespace fk {

// Thread-local last error string returned by fk_last_error().
void set_error(const char* fmt, ...);

#define FK_CHECK_HIP(expr)                                                       \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess) {                                                      \
      ::fk::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),    \
                      __FILE__, __LINE__);                                       \
      return FK_ERR_HIP;                                                         \
    }                                                                            \
  } while (0)

#define FK_REQUIRE(cond, ...)                                                    \
  do {                                                                           \
    if (!(cond)) {                                                               \
      ::fk::set_error(__VA_ARGS__);                                              \
      return FK_ERR_ARG;                                                         \
    }                                                                            \
  } while (0)

// Launch check: kernel launches report configuration errors asynchronously via
// hipGetLastError().
#define FK_CHECK_LAUNCH() FK_CHECK_HIP(hipGetLastError())

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

inline int div_up(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

int device_cu_count();

}  // namespace fk

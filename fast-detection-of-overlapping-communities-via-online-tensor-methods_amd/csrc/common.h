// Internal helpers shared by the HIP kernels and the host pipeline of the MMSB
// tensor-method library: status codes, error recording, launch checks and the
// MFMA vector types used on gfx950 (CDNA4, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "mmsb_stpm.h"

namespace mmsb {

// Records a message for mmsb_last_error() (thread-local) and returns `code`.
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

#define MMSB_HIP(expr)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return ::mmsb::set_error(MMSB_ERR_HIP, "%s:%d %s -> %s", __FILE__, __LINE__,      \
                               #expr, hipGetErrorString(e_));                           \
  } while (0)

#define MMSB_CHECK_ARG(cond, ...)                                                       \
  do {                                                                                  \
    if (!(cond)) return ::mmsb::set_error(MMSB_ERR_ARG, __VA_ARGS__);                   \
  } while (0)

#define MMSB_TRY(expr)                                                                  \
  do {                                                                                  \
    int s_ = (expr);                                                                    \
    if (s_ != MMSB_OK) return s_;                                                       \
  } while (0)

// Launch error check (kernel launches are asynchronous; this catches bad configs).
#define MMSB_LAUNCH_CHECK() MMSB_HIP(hipGetLastError())

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t round_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

}  // namespace mmsb

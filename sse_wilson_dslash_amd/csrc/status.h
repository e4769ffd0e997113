// Per-thread error state of the C ABI and argument checks shared by the
// translation units of libdslash_mi355x.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include "dslash_mi355x.h"

namespace wd {

int set_error(int code, const char* fmt, ...);
void clear_error();

}  // namespace wd

#define WD_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return wd::set_error(DSLASH_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                           __FILE__, __LINE__);                                       \
  } while (0)

#define WD_LAUNCHED()                                                                   \
  do {                                                                                  \
    hipError_t e_ = hipGetLastError();                                                  \
    if (e_ != hipSuccess)                                                               \
      return wd::set_error(DSLASH_ERR_HIP, "kernel launch: %s (%s:%d)", hipGetErrorString(e_), \
                           __FILE__, __LINE__);                                         \
  } while (0)

#define WD_REQUIRE(cond, ...)                                        \
  do {                                                               \
    if (!(cond)) return wd::set_error(DSLASH_ERR_INVALID, __VA_ARGS__); \
  } while (0)

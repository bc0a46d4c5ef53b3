// lj_util.h -- status/error helpers and device allocation of libljhip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "lj_hip.h"

namespace lj {

// records the message for lj_last_error() on this thread and returns code
int fail(int code, const char* fmt, ...);
const std::string& last_error();

#define LJ_HIP(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return ::lj::fail(LJ_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                \
  } while (0)

#define LJ_TRY(expr)              \
  do {                            \
    int rc_ = (expr);             \
    if (rc_ != LJ_OK) return rc_; \
  } while (0)

template <class T>
int dev_alloc(T** p, size_t count)
{
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(LJ_ERR_ALLOC, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
  }
  return LJ_OK;
}

template <class T>
void dev_free(T*& p)
{
  if (p) (void)hipFree(p);
  p = nullptr;
}

inline int check_device()
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(LJ_ERR_DEVICE, "no HIP device available (the HIP path has no CPU fallback)");
  return LJ_OK;
}

}  // namespace lj

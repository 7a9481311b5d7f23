// common.hpp -- shared host/device helpers of the gfx950 FilSim library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/filsim.h"

namespace fs {

// Last error message, per host thread.
void set_error(const char *fmt, ...);

#define FS_HIP(call)                                                       \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess) {                                                \
      fs::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                   \
      return FS_ERR_HIP;                                                   \
    }                                                                      \
  } while (0)

#define FS_CHECK(cond, code, ...)  \
  do {                             \
    if (!(cond)) {                 \
      fs::set_error(__VA_ARGS__);  \
      return (code);               \
    }                              \
  } while (0)

#define FS_TRY(expr)         \
  do {                       \
    int rc_ = (expr);        \
    if (rc_ != FS_OK) return rc_; \
  } while (0)

// Kernel-launch check: catches invalid launch configurations immediately.
#define FS_LAUNCHED()                                                      \
  do {                                                                     \
    hipError_t e_ = hipGetLastError();                                     \
    if (e_ != hipSuccess) {                                                \
      fs::set_error("kernel launch failed: %s (%s:%d)",                    \
                    hipGetErrorString(e_), __FILE__, __LINE__);            \
      return FS_ERR_HIP;                                                   \
    }                                                                      \
  } while (0)

// Complex value, aligned to its size so that LDS and global accesses of one
// value are single 8-byte (FP32) or 16-byte (FP64) instructions.
template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) {
  return {a.x + b.x, a.y + b.y};
}
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) {
  return {a.x - b.x, a.y - b.y};
}

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// Device buffer owned by a handle.
struct dbuf {
  void *p = nullptr;
  size_t bytes = 0;
  int alloc(size_t n) {
    if (n <= bytes && p) return FS_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (n == 0) return FS_OK;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      set_error("hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
      p = nullptr;
      return FS_ERR_HIP;
    }
    bytes = n;
    return FS_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

}  // namespace fs

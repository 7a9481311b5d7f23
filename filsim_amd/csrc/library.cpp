// library.cpp -- library-wide C-ABI: errors, version, device count.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "common.hpp"

namespace fs {
static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}
}  // namespace fs

extern "C" const char *fs_last_error(void) { return fs::g_err; }

extern "C" int fs_version(void) { return 10000; }

extern "C" int fs_device_count(int *count) {
  FS_CHECK(count, FS_ERR_ARG, "null argument");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return FS_OK;
}

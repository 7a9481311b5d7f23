// warmstart.hpp -- optional learned warm start attached to a system.
#pragma once

#include "system_impl.hpp"

namespace fs {

struct WarmStart {
  bool attached() const { return false; }
  int timestep(SystemImpl &, int, int *) { return FS_ERR_STATE; }
  void release() {}
};

}  // namespace fs

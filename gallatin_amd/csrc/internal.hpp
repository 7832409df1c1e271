// Host-side handle of one allocator instance (private to libgallatin_hip).
#ifndef GALLATIN_CAPI_INTERNAL_HPP
#define GALLATIN_CAPI_INTERNAL_HPP

#include <hip/hip_runtime.h>

#include <string>

#include "gallatin/allocators/layout.hpp"
#include "gallatin_capi.h"

struct gal_allocator {
  gallatin::allocators::gal_state host;  // host copy of the device object
  gallatin::allocators::gal_state* dev;  // device object (inside `meta`)
  void* meta = nullptr;                  // metadata arena
  uint64_t meta_bytes = 0;
  void* pool = nullptr;
  uint64_t pool_bytes = 0;
  int pool_kind = 0;                     // 0 device, 1 pinned host, 2 managed
  uint64_t* tree_base = nullptr;         // device address of the tree arena
  uint64_t tree_words = 0;
  int device = 0;
};

namespace gallatin_capi {
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// Workgroups of `threads` for n threads, and the status of the last launch.
inline unsigned grid_for(uint64_t n, uint32_t threads) {
  return (unsigned)((n + threads - 1) / threads);
}
inline int launch_status(const char* what) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GAL_OK : hip_fail(e, what);
}
}  // namespace gallatin_capi

#endif

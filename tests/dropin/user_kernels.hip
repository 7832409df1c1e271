// A user's translation unit written against the allocator's C++ headers the
// way programs written for the reference use it: device malloc/free inlined
// into the user's own kernels, its own global allocator symbol, and a
// Gallatin<> instance it creates and frees itself.  Reference usage followed:
// tests/src/global_test.cu:33-120 (insert_one_size, free_one_size_pointer)
// and tests/src/gallatin_test.cu:59-112.
//
// Built into tests/dropin/libuser_kernels.so (linked with libgallatin_hip.so)
// and driven from tests/test_gpu_dropin.py.
#include <hip/hip_runtime.h>

#include "gallatin/allocators/global_allocator.hpp"

using namespace gallatin::allocators;

namespace {

__global__ void insert_one_size(uint64_t num_inserts, uint64_t size, uint64_t** ptrs,
                                unsigned long long* misses) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid >= num_inserts) return;
  uint64_t* p = static_cast<uint64_t*>(global_malloc(size));
  ptrs[tid] = p;
  if (p == nullptr) {
    atomicAdd(misses, 1ULL);
    return;
  }
  p[0] = tid;
}

__global__ void free_one_size(uint64_t num_inserts, uint64_t** ptrs,
                              unsigned long long* bad) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid >= num_inserts) return;
  uint64_t* p = ptrs[tid];
  if (p == nullptr) return;
  if (p[0] != tid) atomicAdd(bad, 1ULL);
  global_free(p);
}

using my_allocator = Gallatin<16ULL * 1024 * 1024, 16ULL, 4096ULL>;

// alloc_one_size_pointer + free in one kernel on an instance the user owns.
__global__ void own_instance_round(my_allocator* alloc, uint64_t n, uint64_t size,
                                   unsigned long long* counts) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid >= n) return;
  uint32_t* p = static_cast<uint32_t*>(alloc->malloc(size));
  if (p == nullptr) {
    atomicAdd(&counts[0], 1ULL);
    return;
  }
  if (!alloc->owns_allocation(p)) atomicAdd(&counts[1], 1ULL);
  if (alloc->get_allocated_size(size) < size) atomicAdd(&counts[1], 1ULL);
  p[0] = (uint32_t)tid;
  if (p[0] != (uint32_t)tid) atomicAdd(&counts[1], 1ULL);
  alloc->free(p);
}

}  // namespace

// counts_out[0] misses, [1] pattern / ownership errors, for each part:
// [0..1] global allocator, [2..3] own instance.
extern "C" int user_run(uint64_t pool_bytes, uint64_t n, uint64_t size,
                        unsigned long long* counts_out) {
  init_global_allocator(pool_bytes, 42, false);
  uint64_t** ptrs = nullptr;
  unsigned long long* counts = nullptr;
  if (hipMalloc(&ptrs, n * sizeof(uint64_t*)) != hipSuccess) return -1;
  if (hipMalloc(&counts, 4 * sizeof(unsigned long long)) != hipSuccess) return -1;
  if (hipMemset(counts, 0, 4 * sizeof(unsigned long long)) != hipSuccess) return -1;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  insert_one_size<<<blocks, 256>>>(n, size, ptrs, counts);
  free_one_size<<<blocks, 256>>>(n, ptrs, counts + 1);
  if (hipDeviceSynchronize() != hipSuccess) return -2;
  print_global_stats();
  free_global_allocator();

  my_allocator* alloc = my_allocator::generate_on_device(pool_bytes, 7, false);
  if (alloc == nullptr) return -3;
  for (int r = 0; r < 4; ++r) own_instance_round<<<blocks, 256>>>(alloc, n, size, counts + 2);
  if (hipDeviceSynchronize() != hipSuccess) return -2;
  alloc->print_info();
  my_allocator::free_on_device(alloc);

  const hipError_t e = hipMemcpy(counts_out, counts, 4 * sizeof(unsigned long long),
                                 hipMemcpyDeviceToHost);
  (void)hipFree(ptrs);
  (void)hipFree(counts);
  return e == hipSuccess ? 0 : -2;
}

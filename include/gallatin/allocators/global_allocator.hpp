// Process-global allocator instance (global_allocator.cuh:41-97).
//
// As in the reference, the device symbol `global_gallatin` is defined in this
// header, so include it from one translation unit per program (or per shared
// library).  init_global_allocator creates the instance through the C ABI
// and stores its device address in the symbol of the including unit.
#ifndef GALLATIN_GLOBAL_ALLOCATOR_HPP
#define GALLATIN_GLOBAL_ALLOCATOR_HPP

#include <hip/hip_runtime.h>
#include <stdio.h>

#include "gallatin.hpp"

namespace gallatin {
namespace allocators {

// global_allocator.cuh:41
using global_allocator_type = Gallatin<16ULL * 1024 * 1024, 16ULL, 4096ULL>;

// global_allocator.cuh:43
__device__ global_allocator_type* global_gallatin;

// global_allocator.cuh:48-56
__host__ inline void init_global_allocator(uint64_t num_bytes, uint64_t seed,
                                           bool print_info = true,
                                           bool running_calloc = false) {
  global_allocator_type* local_copy = global_allocator_type::generate_on_device(
      num_bytes, seed, print_info, running_calloc);
  if (hipMemcpyToSymbol(HIP_SYMBOL(global_gallatin), &local_copy,
                        sizeof(local_copy)) != hipSuccess)
    fprintf(stderr, "Gallatin: could not publish the global allocator\n");
  (void)hipDeviceSynchronize();
}

// global_allocator.cuh:59-70
__host__ inline void free_global_allocator() {
  global_allocator_type* local_copy = nullptr;
  if (hipMemcpyFromSymbol(&local_copy, HIP_SYMBOL(global_gallatin),
                          sizeof(local_copy)) != hipSuccess)
    return;
  (void)hipDeviceSynchronize();
  if (local_copy) global_allocator_type::free_on_device(local_copy);
  global_allocator_type* null_copy = nullptr;
  (void)hipMemcpyToSymbol(HIP_SYMBOL(global_gallatin), &null_copy,
                          sizeof(null_copy));
}

// global_allocator.cuh:72-82
__device__ inline void* global_malloc(uint64_t num_bytes) {
  return global_gallatin->malloc(num_bytes);
}
__device__ inline void global_free(void* ptr) { global_gallatin->free(ptr); }

// global_allocator.cuh:86-97
__host__ inline void print_global_stats() {
  global_allocator_type* local_copy = nullptr;
  if (hipMemcpyFromSymbol(&local_copy, HIP_SYMBOL(global_gallatin),
                          sizeof(local_copy)) != hipSuccess)
    return;
  (void)hipDeviceSynchronize();
  if (local_copy) local_copy->print_info();
}

}  // namespace allocators
}  // namespace gallatin

#endif

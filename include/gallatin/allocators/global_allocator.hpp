// Process-global allocator instances (global_allocator.cuh:41-357).
//
// As in the reference, the device symbols `global_gallatin` and
// `global_host_gallatin` are defined in this header, so include it from one
// translation unit per program (or per shared library).
// init_global_allocator creates the instance through the C ABI and stores
// its device address in the symbol of the including unit; the _host
// functions do the same for a second instance whose pool is pinned host
// memory.  The combined and fused functions use both instances: fused
// malloc spills to the host pool when the device pool has no room, and
// fused free sends a pointer to the instance that owns it.  The poison
// functions surround an allocation with size records and a byte pattern
// and check them on free.
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

// ---- the host-pool instance (global_allocator.cuh:44, 100-136) ----

__device__ global_allocator_type* global_host_gallatin;

__host__ inline void init_global_allocator_host(uint64_t num_bytes, uint64_t seed,
                                                bool print_info = true,
                                                bool running_calloc = false) {
  global_allocator_type* local_copy = global_allocator_type::generate_on_device_host(
      num_bytes, seed, print_info, running_calloc);
  if (hipMemcpyToSymbol(HIP_SYMBOL(global_host_gallatin), &local_copy,
                        sizeof(local_copy)) != hipSuccess)
    fprintf(stderr, "Gallatin: could not publish the host allocator\n");
  (void)hipDeviceSynchronize();
}

__host__ inline void free_global_allocator_host() {
  global_allocator_type* local_copy = nullptr;
  if (hipMemcpyFromSymbol(&local_copy, HIP_SYMBOL(global_host_gallatin),
                          sizeof(local_copy)) != hipSuccess)
    return;
  (void)hipDeviceSynchronize();
  if (local_copy) global_allocator_type::free_on_device(local_copy);
  global_allocator_type* null_copy = nullptr;
  (void)hipMemcpyToSymbol(HIP_SYMBOL(global_host_gallatin), &null_copy,
                          sizeof(null_copy));
}

__device__ inline void* global_malloc_host(uint64_t num_bytes) {
  return global_host_gallatin->malloc(num_bytes);
}
__device__ inline void global_free_host(void* ptr) { global_host_gallatin->free(ptr); }

__host__ inline void print_global_stats_host() {
  global_allocator_type* local_copy = nullptr;
  if (hipMemcpyFromSymbol(&local_copy, HIP_SYMBOL(global_host_gallatin),
                          sizeof(local_copy)) != hipSuccess)
    return;
  (void)hipDeviceSynchronize();
  if (local_copy) local_copy->print_info();
}

// ---- both instances (global_allocator.cuh:140-224) ----

__host__ inline void init_global_allocator_combined(uint64_t num_bytes, uint64_t host_bytes,
                                                    uint64_t seed, bool print_info = true,
                                                    bool running_calloc = false) {
  init_global_allocator(num_bytes, seed, print_info, running_calloc);
  init_global_allocator_host(host_bytes, seed, print_info, running_calloc);
}

__host__ inline void free_global_allocator_combined() {
  free_global_allocator();
  free_global_allocator_host();
}

__device__ inline void* global_malloc_combined(uint64_t num_bytes, bool on_host = false) {
  return on_host ? global_malloc_host(num_bytes) : global_malloc(num_bytes);
}

__device__ inline void global_free_combined(void* ptr, bool on_host = false) {
  if (on_host)
    global_free_host(ptr);
  else
    global_free(ptr);
}

// global_allocator.cuh:195-204: device pool first, the host pool when it
// has no room.  Pointers stay where they were placed until freed.
__device__ inline void* global_malloc_fused(uint64_t num_bytes) {
  void* alloc = global_malloc(num_bytes);
  return alloc ? alloc : global_malloc_host(num_bytes);
}

// global_allocator.cuh:207-216.
__device__ inline void global_free_fused(void* ptr) {
  if (global_gallatin->owns_allocation(ptr))
    global_free(ptr);
  else
    global_free_host(ptr);
}

__host__ inline void print_global_stats_combined() {
  printf("Device Allocator:\n");
  print_global_stats();
  printf("Host Allocator:\n");
  print_global_stats_host();
}

// ---- poison (global_allocator.cuh:238-357) ----
//
// Layout of a poisoned allocation of n bytes inside a slot of n + 32 + x
// bytes (x = the slot's padding): [n, x] (16 B), the caller's n bytes, x
// bytes of pattern (byte i = (char)i), [n, x] (16 B).  The caller gets the
// address after the first record.

__device__ inline void* global_malloc_poison(uint64_t bytes_needed) {
  const uint64_t extra_bytes =
      global_gallatin->get_allocated_size(bytes_needed + 32) - (bytes_needed + 32);
  void* alloced_memory = global_gallatin->malloc(bytes_needed + 32);
  if (!alloced_memory) return nullptr;
  const uint64_t base = reinterpret_cast<uint64_t>(alloced_memory);
  uint64_t* poison_start = reinterpret_cast<uint64_t*>(base);
  uint64_t* poison_end = reinterpret_cast<uint64_t*>(base + 16 + bytes_needed + extra_bytes);
  char* memory_end = reinterpret_cast<char*>(base + 16 + bytes_needed);
  for (uint64_t i = 0; i < extra_bytes; ++i) memory_end[i] = (char)i;
  utils::atomic_store(&poison_start[0], bytes_needed);
  utils::atomic_store(&poison_start[1], extra_bytes);
  utils::atomic_store(&poison_end[0], bytes_needed);
  utils::atomic_store(&poison_end[1], extra_bytes);
  return reinterpret_cast<void*>(base + 16);
}

// True when both size records and the padding pattern are intact; prints
// which check failed otherwise, as the reference does.
__device__ inline bool global_check_poison(void* allocation) {
  if (allocation == nullptr) {
    printf("Poison violated 0\n");
    return false;
  }
  const uint64_t alloc_as_bytes = reinterpret_cast<uint64_t>(allocation);
  const uint64_t* poison_start = reinterpret_cast<const uint64_t*>(alloc_as_bytes - 16);
  const uint64_t bytes_needed = poison_start[0];
  const uint64_t extra_bytes = poison_start[1];
  if ((bytes_needed + extra_bytes) % 16 != 0) {
    printf("Poison violated 1\n");
    return false;
  }
  const char* check_region = reinterpret_cast<const char*>(alloc_as_bytes + bytes_needed);
  const uint64_t* poison_end =
      reinterpret_cast<const uint64_t*>(alloc_as_bytes + bytes_needed + extra_bytes);
  if (poison_end[0] != bytes_needed) {
    printf("Poison violated 2\n");
    return false;
  }
  if (poison_end[1] != extra_bytes) {
    printf("Poison violated 3\n");
    return false;
  }
  for (uint64_t i = 0; i < extra_bytes; ++i) {
    if (check_region[i] != (char)i) {
      printf("Poison 4 at %llu bytes after allocation\n", (unsigned long long)extra_bytes);
      return false;
    }
  }
  return true;
}

__device__ inline void global_free_poison(void* allocation) {
  global_check_poison(allocation);
  if (!allocation) return;  // the reference would free address -16
  global_free(reinterpret_cast<void*>(reinterpret_cast<uint64_t>(allocation) - 16));
}

}  // namespace allocators
}  // namespace gallatin

#endif

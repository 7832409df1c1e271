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
// A poisoned allocation of n bytes takes the slot a request of n + 32 bytes
// gets, of S = get_allocated_size(n + 32) bytes, laid out as
//
//   [0, 16)                 head record {n, x}
//   [16, 16 + n)            the caller's bytes (the pointer returned)
//   [16 + n, 16 + n + x)    x = S - n - 32 bytes of padding, byte i = (char)i
//   [16 + n + x, S)         tail record {n, x}
//
// The records and the pattern are the reference's observable layout; the
// check reports the first broken part with the reference's messages.

struct poison_record {
  uint64_t size;     // n, the bytes the caller asked for
  uint64_t padding;  // x, the pattern bytes between the caller's and the tail
};

__device__ inline poison_record* poison_head(void* user) {
  return reinterpret_cast<poison_record*>(static_cast<char*>(user) - sizeof(poison_record));
}

__device__ inline poison_record* poison_tail(void* user, const poison_record& r) {
  return reinterpret_cast<poison_record*>(static_cast<char*>(user) + r.size + r.padding);
}

// Records are written with device-coherent stores, so a checker in another
// workgroup or kernel sees them.
__device__ inline void poison_store(poison_record* at, const poison_record& r) {
  utils::atomic_store(&at->size, r.size);
  utils::atomic_store(&at->padding, r.padding);
}

// A failed malloc returns nullptr (the reference writes its records through
// the null pointer).
__device__ inline void* global_malloc_poison(uint64_t bytes_needed) {
  constexpr uint64_t kRecords = 2 * sizeof(poison_record);
  const uint64_t slot = global_gallatin->get_allocated_size(bytes_needed + kRecords);
  char* raw = static_cast<char*>(global_gallatin->malloc(bytes_needed + kRecords));
  if (raw == nullptr) return nullptr;
  const poison_record rec{bytes_needed, slot - bytes_needed - kRecords};
  char* user = raw + sizeof(poison_record);
  char* pad = user + rec.size;
  for (uint64_t i = 0; i < rec.padding; ++i) pad[i] = (char)i;
  poison_store(poison_head(user), rec);
  poison_store(poison_tail(user, rec), rec);
  return user;
}

// True when both records and the padding pattern are intact.  The checks
// run in the reference's order: null pointer (0), head record shape (1),
// tail size (2), tail padding (3, where the reference's bool function has a
// bare return), pattern (4).
__device__ inline bool global_check_poison(void* allocation) {
  if (allocation == nullptr) {
    printf("Poison violated 0\n");
    return false;
  }
  const poison_record head = *poison_head(allocation);
  if ((head.size + head.padding) % 16 != 0) {
    printf("Poison violated 1\n");
    return false;
  }
  const poison_record tail = *poison_tail(allocation, head);
  if (tail.size != head.size) {
    printf("Poison violated 2\n");
    return false;
  }
  if (tail.padding != head.padding) {
    printf("Poison violated 3\n");
    return false;
  }
  const char* pad = static_cast<const char*>(allocation) + head.size;
  for (uint64_t i = 0; i < head.padding; ++i) {
    if (pad[i] != (char)i) {
      printf("Poison 4 at %llu bytes after allocation\n", (unsigned long long)head.padding);
      return false;
    }
  }
  return true;
}

// Checks, then frees the whole slot; nullptr is checked (and reported) but
// not freed (the reference would free address -16).
__device__ inline void global_free_poison(void* allocation) {
  global_check_poison(allocation);
  if (allocation == nullptr) return;
  global_free(poison_head(allocation));
}

}  // namespace allocators
}  // namespace gallatin

#endif

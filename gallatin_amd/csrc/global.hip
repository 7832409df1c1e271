// The library's global allocator instance and its bulk kernels
// (global_allocator.cuh:41-97).  This is the only translation unit of the
// library that includes global_allocator.hpp, which defines the device
// symbol global_gallatin.
#include <hip/hip_runtime.h>

#include "gallatin/allocators/global_allocator.hpp"
#include "gallatin_capi.h"
#include "internal.hpp"

using namespace gallatin::allocators;
using namespace gallatin_capi;

__global__ void k_global_malloc_bulk(const uint64_t* sizes, uint64_t n, void** out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = global_malloc(sizes[i]);
}

__global__ void k_global_free_bulk(void* const* ptrs, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (ptrs[i]) global_free(ptrs[i]);
}

static global_allocator_type* current_global() {
  global_allocator_type* p = nullptr;
  if (hipMemcpyFromSymbol(&p, HIP_SYMBOL(global_gallatin), sizeof(p)) != hipSuccess)
    return nullptr;
  return p;
}

extern "C" {

int gal_init_global(uint64_t num_bytes, uint64_t seed, int print_info) {
  if (current_global()) return fail(GAL_ERR_INVALID, "global allocator already initialised");
  init_global_allocator(num_bytes, seed, print_info != 0, false);
  if (!current_global()) return fail(GAL_ERR_BOOT, "init_global_allocator failed");
  return GAL_OK;
}

int gal_free_global(void) {
  if (!current_global()) return fail(GAL_ERR_INVALID, "no global allocator");
  free_global_allocator();
  return GAL_OK;
}

int gal_print_global_stats(void) {
  if (!current_global()) return fail(GAL_ERR_INVALID, "no global allocator");
  print_global_stats();
  return GAL_OK;
}

gal_allocator* gal_global_handle(void) { return gal_from_device_state(current_global()); }

int gal_global_malloc_bulk(const uint64_t* sizes, uint64_t n, void** out, uint32_t threads,
                           void* stream) {
  if ((!sizes || !out) && n) return fail(GAL_ERR_INVALID, "null argument");
  if (!current_global()) return fail(GAL_ERR_INVALID, "no global allocator");
  if (n == 0) return GAL_OK;
  if (threads == 0 || threads > 1024 || threads % 64) threads = 256;
  k_global_malloc_bulk<<<(unsigned)((n + threads - 1) / threads), threads, 0,
                         (hipStream_t)stream>>>(sizes, n, out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GAL_OK : hip_fail(e, "global_malloc_bulk");
}

int gal_global_free_bulk(void* const* ptrs, uint64_t n, uint32_t threads, void* stream) {
  if (!ptrs && n) return fail(GAL_ERR_INVALID, "null argument");
  if (!current_global()) return fail(GAL_ERR_INVALID, "no global allocator");
  if (n == 0) return GAL_OK;
  if (threads == 0 || threads > 1024 || threads % 64) threads = 256;
  k_global_free_bulk<<<(unsigned)((n + threads - 1) / threads), threads, 0,
                       (hipStream_t)stream>>>(ptrs, n);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GAL_OK : hip_fail(e, "global_free_bulk");
}

}  // extern "C"

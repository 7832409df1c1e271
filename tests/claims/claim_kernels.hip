// Test kernels for the claim-level parity of the batched forms: one
// workgroup batch (batch_malloc, include/gallatin/allocators/batched.hpp)
// and one slice-cache refill (cache_malloc, slice_cache.hpp) on an instance
// the test created, with every request's byte offset written out, so the
// test can compare the slices handed out per class with the oracle's
// coalesced claims (gallatin.cuh:865-1078, block.cuh:245-281).
//
// Built into tests/claims/libclaim_kernels.so (linked with
// libgallatin_hip.so) by tests/claims/build.py; driven from
// tests/test_gpu_claims.py.
#include <hip/hip_runtime.h>

#include "gallatin/allocators/batched.hpp"
#include "gallatin/allocators/slice_cache.hpp"

using namespace gallatin::allocators;

namespace {

template <int kThreads>
__global__ __launch_bounds__(kThreads) void k_claim_batch(gal_state* s, const uint64_t* sizes,
                                                          uint32_t n, uint64_t* out) {
  __shared__ batch::lds_state L;
  batch::batch_init(L);
  const uint32_t i = threadIdx.x;
  const uint64_t size = i < n ? sizes[i] : 0;
  void* p = batch::batch_malloc(s, size != 0, size, L);
  if (i < n) out[i] = p ? (uint64_t)((char*)p - s->memory) : ~0ULL;
}

__global__ __launch_bounds__(64) void k_claim_cache(gal_state* s, const uint64_t* sizes,
                                                    uint32_t n, uint64_t* out) {
  __shared__ cache::slice_cache L;
  cache::cache_init(L);
  const uint32_t i = threadIdx.x;
  const uint64_t size = i < n ? sizes[i] : 0;
  void* p = cache::cache_malloc(s, L, size != 0, size);
  if (i < n) out[i] = p ? (uint64_t)((char*)p - s->memory) : ~0ULL;
  cache::cache_flush(s, L);
}

}  // namespace

// One workgroup of `threads` (256, 512 or 1024) threads; thread i < n asks
// for sizes_dev[i] bytes (0: no request).  out_dev[i]: byte offset from the
// pool base or ~0.  Synchronous; returns 0 or a negative error.
extern "C" int claims_batch(void* dev_state, const uint64_t* sizes_dev, uint32_t n,
                            uint64_t* out_dev, uint32_t threads) {
  gal_state* s = static_cast<gal_state*>(dev_state);
  if (!s || n > threads) return -1;
  if (threads == 256)
    k_claim_batch<256><<<1, 256>>>(s, sizes_dev, n, out_dev);
  else if (threads == 512)
    k_claim_batch<512><<<1, 512>>>(s, sizes_dev, n, out_dev);
  else if (threads == 1024)
    k_claim_batch<1024><<<1, 1024>>>(s, sizes_dev, n, out_dev);
  else
    return -1;
  return hipDeviceSynchronize() == hipSuccess ? 0 : -2;
}

// One wave (64 threads) through a fresh slice cache, then its flush.
extern "C" int claims_cache(void* dev_state, const uint64_t* sizes_dev, uint32_t n,
                            uint64_t* out_dev) {
  gal_state* s = static_cast<gal_state*>(dev_state);
  if (!s || n > 64) return -1;
  k_claim_cache<<<1, 64>>>(s, sizes_dev, n, out_dev);
  return hipDeviceSynchronize() == hipSuccess ? 0 : -2;
}

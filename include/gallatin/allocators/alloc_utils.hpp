// Integer helpers, agent-scope atomics and wave64 primitives used by the
// device allocator.  gfx950 only: wavefronts are 64 lanes, ballots are 64-bit.
//
// Reference counterparts: alloc_utils.cuh:382-398 (get_first_bit_bigger),
// alloc_utils.cuh:440-453 (__cfcll), alloc_utils.cuh:41-147 (cache-hinted
// loads, replaced here by agent-scope relaxed atomic loads so that every XCD
// sees the same metadata), and the cooperative-groups coalesced_group /
// labeled_partition / exclusive_scan / reduce / ballot calls of
// gallatin.cuh:883-1033 and block.cuh:245-281 (replaced by ballot/mbcnt code).
#ifndef GALLATIN_ALLOC_UTILS_HPP
#define GALLATIN_ALLOC_UTILS_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gallatin {
namespace utils {

// ceil(log2(x)) for x >= 1 (alloc_utils.cuh:382-398).
__host__ __device__ inline int get_first_bit_bigger(uint64_t x) {
  return 63 - __builtin_clzll(x) + (__builtin_popcountll(x) != 1);
}

// Number of trailing one bits (alloc_utils.cuh:440-453).
__host__ __device__ inline int cfcll(uint64_t bits) {
  return bits == ~0ULL ? 64 : __builtin_ctzll(~bits);
}

__host__ __device__ inline uint64_t low_mask(uint64_t n) {
  return n >= 64 ? ~0ULL : ((1ULL << n) - 1ULL);
}

// ---------------------------------------------------------------- atomics
// All shared metadata is accessed with agent-scope atomics: gfx950 has one
// L2 per XCD, and only atomics / sc1 loads are coherent across them.

template <class T>
__device__ __forceinline__ T atomic_load(const T* p) {
  return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void atomic_store(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T atomic_add(T* p, T v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T atomic_sub(T* p, T v) {
  return __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T atomic_or(T* p, T v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T atomic_and(T* p, T v) {
  return __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T atomic_max(T* p, T v) {
  return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T atomic_exch(T* p, T v) {
  return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Returns the value found (== cmp on success).
template <class T>
__device__ __forceinline__ T atomic_cas(T* p, T cmp, T v) {
  __hip_atomic_compare_exchange_strong(p, &cmp, v, __ATOMIC_RELAXED,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return cmp;
}

// Ordering of payload and metadata across XCDs.
__device__ __forceinline__ void fence_release() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
}
__device__ __forceinline__ void fence_acquire() {
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}
__device__ __forceinline__ void fence_acq_rel() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
}

__device__ __forceinline__ void backoff() { __builtin_amdgcn_s_sleep(2); }

// Global thread index of a 1-D launch (alloc_utils.cuh:400-402).
__device__ __forceinline__ uint64_t get_tid() {
  return (uint64_t)threadIdx.x + (uint64_t)blockIdx.x * (uint64_t)blockDim.x;
}

}  // namespace utils

// ------------------------------------------------------------ wave64 ops
namespace wave {

__device__ __forceinline__ uint64_t active() { return __ballot(1); }
__device__ __forceinline__ uint32_t lane() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanes_below() {
  return (1ULL << __lane_id()) - 1ULL;
}
__device__ __forceinline__ int first(uint64_t m) { return __builtin_ctzll(m); }
__device__ __forceinline__ uint32_t count(uint64_t m) {
  return (uint32_t)__builtin_popcountll(m);
}

// Value of `v` in lane `l` (l must be wave-uniform).
__device__ __forceinline__ uint32_t readlane(uint32_t v, int l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ uint64_t readlane(uint64_t v, int l) {
  uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t readfirst(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ uint64_t readfirst(uint64_t v) {
  uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}
// Value of `v` in lane `src` (per-lane source, any active lane).
__device__ __forceinline__ uint32_t fetch(uint32_t v, int src) {
  return (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)v);
}
__device__ __forceinline__ uint64_t fetch(uint64_t v, int src) {
  uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)v);
  uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// True on the lowest active lane.
__device__ __forceinline__ bool is_first() {
  return __lane_id() == (uint32_t)__builtin_ctzll(__ballot(1));
}

// Partition the lanes of `cand` (a ballot of `in`) by `key`: each lane with
// `in` set gets the mask of lanes sharing its key; others get 0.  Replaces
// cooperative_groups::labeled_partition (gallatin.cuh:883-885, 1784-1786).
// One ballot per distinct key; must be reached by every active lane.
__device__ __forceinline__ uint64_t match(uint64_t cand, bool in, uint64_t key) {
  uint64_t mine = 0;
  uint64_t rem = cand;
  while (rem) {
    int l = first(rem);
    uint64_t k = readlane(key, l);
    bool eq = in && key == k;
    uint64_t g = __ballot(eq);
    if (eq) mine = g;
    rem &= ~g;
  }
  return mine;
}

// Sum of `v` over the lanes of `group` that lie below this lane (exclusive
// prefix) and over the whole group, from bit-sliced ballots.  Exact for any
// exec mask; `max_bits` bounds the values.  Replaces cg::exclusive_scan and
// cg::reduce over a coalesced group (block.cuh:245-281).
__device__ __forceinline__ void group_scan(uint32_t v, uint64_t group,
                                           uint32_t& prefix, uint32_t& total) {
  uint32_t p = 0, t = 0;
  const uint64_t below = lanes_below();
  uint64_t any = __ballot(v != 0);
  for (int b = 0; any; ++b) {
    uint64_t bits = __ballot((v >> b) & 1u);
    p += count(bits & group & below) << b;
    t += count(bits & group) << b;
    any = (b + 1 < 32) ? __ballot((v >> (b + 1)) != 0) : 0;
  }
  prefix = p;
  total = t;
}

}  // namespace wave
}  // namespace gallatin

#endif

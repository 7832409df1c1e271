// MurmurHash64A (Austin Appleby's MurmurHash2, 64-bit "A" variant), as used
// by the reference for the hashed segment start (murmurhash.cuh:20-61,
// veb.cuh:946-977).  Written from the published algorithm.
#ifndef GALLATIN_MURMURHASH_HPP
#define GALLATIN_MURMURHASH_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gallatin {
namespace hashers {

// Hash of one 64-bit key (len = 8), the only form the allocator needs.
__host__ __device__ inline uint64_t MurmurHash64A_u64(uint64_t key,
                                                      uint64_t seed) {
  const uint64_t m = 0xc6a4a7935bd1e995ULL;
  const int r = 47;
  uint64_t h = seed ^ (8ULL * m);
  uint64_t k = key;
  k *= m;
  k ^= k >> r;
  k *= m;
  h ^= k;
  h *= m;
  h ^= h >> r;
  h *= m;
  h ^= h >> r;
  return h;
}

}  // namespace hashers
}  // namespace gallatin

#endif

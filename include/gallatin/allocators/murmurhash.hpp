// MurmurHash64A (Austin Appleby's MurmurHash2, 64-bit "A" variant), as used
// by the reference for the hashed segment start (murmurhash.cuh:20-61,
// veb.cuh:946-977).  Written from the published algorithm.
#ifndef GALLATIN_MURMURHASH_HPP
#define GALLATIN_MURMURHASH_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gallatin {
namespace hashers {

// Hash of one 64-bit key (len = 8), the form the allocator needs.
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

// Hash of `len` bytes at `key` (any alignment), the general form the
// reference's hash tables use for their keys.
__host__ __device__ inline uint64_t MurmurHash64A(const void* key, int len, uint64_t seed) {
  const uint64_t m = 0xc6a4a7935bd1e995ULL;
  const int r = 47;
  const unsigned char* data = static_cast<const unsigned char*>(key);
  uint64_t h = seed ^ ((uint64_t)len * m);
  const int blocks = len / 8;
  for (int b = 0; b < blocks; ++b) {
    uint64_t k = 0;
    for (int i = 7; i >= 0; --i) k = (k << 8) | data[8 * b + i];  // little-endian word
    k *= m;
    k ^= k >> r;
    k *= m;
    h ^= k;
    h *= m;
  }
  const unsigned char* tail = data + 8 * blocks;
  const int rest = len & 7;
  if (rest) {
    for (int i = rest - 1; i >= 0; --i) h ^= (uint64_t)tail[i] << (8 * i);
    h *= m;
  }
  h ^= h >> r;
  h *= m;
  h ^= h >> r;
  return h;
}

}  // namespace hashers
}  // namespace gallatin

#endif

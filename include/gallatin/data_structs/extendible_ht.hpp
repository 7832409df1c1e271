// Allocator-backed extendible hash table of 64-bit keys and values
// (extendible_ht.cuh:132-1437).
//
// Layout.  Buckets of items_per_bucket pairs come from the global allocator.
// Bucket b has a home index h(b) and a depth d(b) >= min_bits and owns the
// keys whose hash agrees with h(b) in the low d(b) bits.  The directory has
// 2^max_bits entries (taken from the global allocator and zeroed once);
// entry i holds the bucket whose home is i, or nothing if no bucket with
// that home exists yet.  The owner of a hash is found by looking at the
// entry of its low `level` bits, then of its low level - 1 bits, ... and
// taking the first bucket present (the reference's refinement from the
// global level down, extendible_ht.cuh:842-1115); the entries of four
// depths are loaded together.  Entries [0, 2^min_bits) are filled at
// creation.
//
// Writes (insert, remove) lock the owner bucket, check that it still owns
// the key and work under the lock.  The lock bit and the depth share one
// word, so the atomic that takes the lock also returns the depth.  A full
// bucket is split: a new bucket with home h(b) + 2^d(b) and depth d(b) + 1
// receives the keys whose bit d(b) is set, is published in the directory,
// then the old bucket's depth is raised and the moved keys are cleared from
// it (the reference's promotion, extendible_ht.cuh:1118-1289).  Queries
// take no lock: they read the owner's depth, scan its keys, and when the key
// is missing look again if the depth changed meanwhile, so a key is never
// missed while it moves; a found key's value counts only if the key is still
// in its slot after the value was read (the slot may have been emptied and
// reused).  Every bucket word is read and written with device-coherent
// atomics; a value is written before its key and a moved key before the new
// bucket is published.  Lock attempts per operation are bounded.
//
// Differences from the reference: insert of a present key replaces its
// value; remove() exists (the reference has none, so its behaviour is not
// pinned by it); a table that cannot split further, or cannot get memory,
// makes insert return false instead of trapping.
#ifndef GALLATIN_EXTENDIBLE_HT_HPP
#define GALLATIN_EXTENDIBLE_HT_HPP

#include <hip/hip_runtime.h>

#include "gallatin/allocators/global_allocator.hpp"
#include "gallatin/allocators/murmurhash.hpp"

namespace gallatin {
namespace data_structs {

template <typename table>
__global__ void ext_ht_alloc_directory(table* t) {
  if (blockIdx.x == 0 && threadIdx.x == 0) t->alloc_directory();
}

template <typename table>
__global__ void ext_ht_first_buckets(table* t, uint64_t n) {
  const uint64_t i = gallatin::utils::get_tid();
  if (i < n) t->make_first_bucket(i);
}

template <typename table>
__global__ void ext_ht_free_buckets(table* t, uint64_t n) {
  const uint64_t i = gallatin::utils::get_tid();
  if (i < n) t->free_bucket_at(i);
}

template <typename table>
__global__ void ext_ht_free_directory(table* t) {
  if (blockIdx.x == 0 && threadIdx.x == 0) t->free_directory();
}

// Keys present (out[0]) and buckets (out[1]): a bucket is the one directory
// entry at its home.  One atomic per wave and counter.
template <typename table>
__global__ void ext_ht_count(table* t, uint64_t n, unsigned long long* out) {
  const uint64_t i = gallatin::utils::get_tid();
  uint64_t keys = 0, buckets = 0;
  if (i < n) {
    keys = t->bucket_fill(i);
    buckets = t->directory && t->directory[i] ? 1 : 0;
  }
  for (int off = 32; off >= 1; off >>= 1) {
    const int src = (int)(gallatin::wave::lane() ^ (uint32_t)off);
    keys += gallatin::wave::fetch(keys, src);
    buckets += gallatin::wave::fetch(buckets, src);
  }
  if (gallatin::wave::lane() == 0) {
    if (keys) atomicAdd(out, (unsigned long long)keys);
    if (buckets) atomicAdd(out + 1, (unsigned long long)buckets);
  }
}

template <typename Key, Key defaultKey, Key tombstoneKey, typename Val, int items_per_bucket,
          uint64_t min_bits, uint64_t max_bits>
struct extendible_hash_table {
  using my_type = extendible_hash_table<Key, defaultKey, tombstoneKey, Val, items_per_bucket,
                                        min_bits, max_bits>;
  static_assert(sizeof(Key) == 8 && sizeof(Val) == 8, "64-bit keys and values");
  static_assert(min_bits >= 1 && min_bits <= max_bits && max_bits <= 40, "table bits");
  static_assert(items_per_bucket >= 2 && items_per_bucket <= 64, "bucket size");

  static constexpr uint64_t min_items = 1ULL << min_bits;
  static constexpr uint64_t max_items = 1ULL << max_bits;
  static constexpr uint32_t kTries = 1u << 22;  // lock attempts per operation

  struct bucket {
    uint64_t ctl;  // (depth << 32) | lock bit
    uint64_t home;
    uint64_t keys[items_per_bucket];
    uint64_t vals[items_per_bucket];
  };

  bucket** directory;  // max_items entries
  uint64_t level;      // deepest bucket depth
  uint64_t seed;
  uint64_t reserved;   // (buckets are counted by count_keys)
  uint64_t failed;     // inserts refused (table full or no memory)

  // ---- host ----
  // extendible_ht.cuh:441-476.  Needs the global allocator.
  static __host__ my_type* generate_on_device(uint64_t seed = 42) {
    my_type host{};
    host.level = min_bits;
    host.seed = seed;
    my_type* dev = nullptr;
    if (hipMalloc(&dev, sizeof(my_type)) != hipSuccess) return nullptr;
    if (hipMemcpy(dev, &host, sizeof(my_type), hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(dev);
      return nullptr;
    }
    ext_ht_alloc_directory<my_type><<<1, 1>>>(dev);
    bucket** dir = nullptr;
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(&dir, &dev->directory, sizeof(dir), hipMemcpyDeviceToHost) != hipSuccess ||
        dir == nullptr || hipMemset(dir, 0, sizeof(bucket*) * max_items) != hipSuccess) {
      (void)hipFree(dev);
      return nullptr;
    }
    ext_ht_first_buckets<my_type><<<(unsigned)((min_items + 255) / 256), 256>>>(dev, min_items);
    if (hipDeviceSynchronize() != hipSuccess) return nullptr;
    return dev;
  }

  // extendible_ht.cuh:479-494.
  static __host__ void free_on_device(my_type* dev) {
    if (!dev) return;
    ext_ht_free_buckets<my_type><<<(unsigned)((max_items + 255) / 256), 256>>>(dev, max_items);
    ext_ht_free_directory<my_type><<<1, 1>>>(dev);
    (void)hipDeviceSynchronize();
    (void)hipFree(dev);
  }

  // Keys present (calculate_fill's count, extendible_ht.cuh:1379-1411) and,
  // into *buckets, the buckets in use.
  static __host__ uint64_t count_keys(my_type* dev, uint64_t* buckets = nullptr) {
    unsigned long long* c = nullptr;
    if (hipMalloc(&c, 2 * sizeof(*c)) != hipSuccess) return 0;
    (void)hipMemset(c, 0, 2 * sizeof(*c));
    ext_ht_count<my_type><<<(unsigned)((max_items + 255) / 256), 256>>>(dev, max_items, c);
    unsigned long long n[2] = {0, 0};
    (void)hipMemcpy(n, c, sizeof(n), hipMemcpyDeviceToHost);
    (void)hipFree(c);
    if (buckets) *buckets = n[1];
    return n[0];
  }

  // ---- device: set-up and teardown ----
  __device__ void alloc_directory() {
    directory = static_cast<bucket**>(
        gallatin::allocators::global_malloc(sizeof(bucket*) * max_items));
  }

  __device__ bucket* new_bucket(uint64_t home, uint32_t depth) {
    bucket* b = static_cast<bucket*>(gallatin::allocators::global_malloc(sizeof(bucket)));
    if (!b) return nullptr;
    // device-coherent stores: another XCD may read the bucket as soon as
    // it is published
    gallatin::utils::atomic_store(&b->ctl, (uint64_t)depth << 32);
    gallatin::utils::atomic_store(&b->home, home);
    for (int i = 0; i < items_per_bucket; ++i)
      gallatin::utils::atomic_store(&b->keys[i], (uint64_t)defaultKey);
    return b;
  }

  __device__ void make_first_bucket(uint64_t i) {
    bucket* b = new_bucket(i, (uint32_t)min_bits);
    if (!b) {
      gallatin::utils::atomic_add(&failed, (uint64_t)1);
      return;
    }
    __builtin_amdgcn_s_waitcnt(0);
    gallatin::utils::atomic_store(reinterpret_cast<uint64_t*>(&directory[i]),
                                  reinterpret_cast<uint64_t>(b));
  }

  __device__ void free_bucket_at(uint64_t i) {
    if (directory && directory[i]) gallatin::allocators::global_free(directory[i]);
  }

  __device__ void free_directory() {
    if (directory) gallatin::allocators::global_free(directory);
    directory = nullptr;
  }

  __device__ uint64_t bucket_fill(uint64_t i) {
    bucket* b = directory ? directory[i] : nullptr;
    if (!b) return 0;
    uint64_t c = 0;
    for (int s = 0; s < items_per_bucket; ++s)
      c += b->keys[s] != defaultKey && b->keys[s] != tombstoneKey;
    return c;
  }

  // ---- device: operations ----
  static __device__ __forceinline__ uint64_t mask(uint64_t d) { return (1ULL << d) - 1; }
  static __device__ __forceinline__ uint64_t ld(const uint64_t* p) {
    return gallatin::utils::atomic_load(p);
  }
  static __device__ __forceinline__ void st(uint64_t* p, uint64_t v) {
    gallatin::utils::atomic_store(p, v);
  }
  // A read that may return an old value: a workgroup-scope load, served by
  // this XCD's caches.  (A volatile read compiles to a system-scope load,
  // which goes to memory like a device-coherent one.)
  static __device__ __forceinline__ uint64_t ld_cached(const uint64_t* p) {
    return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
  }

  __device__ uint64_t hash(Key k) const { return hashers::MurmurHash64A_u64((uint64_t)k, seed); }

  // All keys of a bucket (and its home) are loaded first, so the loads are
  // in flight together, then compared: the slot holding `key` or -1, and
  // the first free slot or -1.
  __device__ uint64_t scan(bucket* b, uint64_t key, int& match, int& empty) {
    uint64_t k[items_per_bucket];
    const uint64_t home = ld(&b->home);
#pragma unroll
    for (int s = 0; s < items_per_bucket; ++s) k[s] = ld(&b->keys[s]);
    match = -1;
    empty = -1;
#pragma unroll
    for (int s = items_per_bucket - 1; s >= 0; --s) {
      if (k[s] == key) match = s;
      if (k[s] == (uint64_t)defaultKey || k[s] == (uint64_t)tombstoneKey) empty = s;
    }
    return home;
  }

  // The bucket owning hash h: the entry of its low `level` bits, or of
  // fewer bits when that home has no bucket yet.  Entries are only ever
  // set, so the deepest one present among the loaded depths is the owner
  // (or a bucket being split, which the callers' home check rejects).
  //
  // `fresh` = false reads `level` through the cache: it may be behind, which
  // only leads to an ancestor of the owner, and every caller checks the
  // bucket's home against its depth and looks again (with fresh = true).
  // Every operation reads `level`, and device-coherent reads of one word
  // from the whole device queue at one memory channel.
  static constexpr int kOwnerFan = 4;
  __device__ bucket* owner(uint64_t h, bool fresh) {
    uint64_t top = fresh ? ld(&level) : ld_cached(&level);
    for (;;) {
      uint64_t e[kOwnerFan];
#pragma unroll
      for (int j = 0; j < kOwnerFan; ++j) {
        const uint64_t d = top >= min_bits + j ? top - j : min_bits;
        e[j] = ld(reinterpret_cast<uint64_t*>(&directory[h & mask(d)]));
      }
#pragma unroll
      for (int j = 0; j < kOwnerFan; ++j)
        if (e[j]) return reinterpret_cast<bucket*>(e[j]);
      if (top < min_bits + kOwnerFan) return nullptr;
      top -= kOwnerFan;
    }
  }

  static __device__ __forceinline__ uint32_t depth_of(uint64_t ctl) {
    return (uint32_t)(ctl >> 32);
  }

  // One attempt; on success `d` is the bucket's depth.  Lock users retry in
  // a loop whose every iteration holds the whole critical section: a lane
  // that gets the lock releases it in the same iteration, so lanes of one
  // wave waiting for the same bucket never keep its holder from running (a
  // spin-then-work loop would).
  __device__ bool try_lock(bucket* b, uint32_t& d) {
    const uint64_t old = gallatin::utils::atomic_or(&b->ctl, (uint64_t)1);
    d = depth_of(old);
    return (old & 1) == 0;
  }

  __device__ void unlock(bucket* b, uint32_t d) {
    __builtin_amdgcn_s_waitcnt(0);  // the critical section's stores land first
    gallatin::utils::atomic_store(&b->ctl, (uint64_t)d << 32);
  }

  // Split full bucket b (locked, depth d): keys with bit d set move to a new
  // bucket with home h(b) + 2^d.  False when no bucket could be had.
  __device__ bool split(bucket* b, uint32_t d, uint64_t old_home) {
    uint64_t k[items_per_bucket], v[items_per_bucket];
#pragma unroll
    for (int s = 0; s < items_per_bucket; ++s) {
      k[s] = ld(&b->keys[s]);
      v[s] = ld(&b->vals[s]);
    }
    const uint64_t home = old_home | (1ULL << d);
    bucket* nb = static_cast<bucket*>(gallatin::allocators::global_malloc(sizeof(bucket)));
    if (!nb) return false;
    uint64_t moved = 0;
#pragma unroll
    for (int s = 0; s < items_per_bucket; ++s) {
      const bool live = k[s] != (uint64_t)defaultKey && k[s] != (uint64_t)tombstoneKey;
      const bool go = live && ((hash((Key)k[s]) >> d) & 1);
      moved |= (uint64_t)go << s;
      if (go) st(&nb->vals[s], v[s]);
      st(&nb->keys[s], go ? k[s] : (uint64_t)defaultKey);
    }
    st(&nb->ctl, (uint64_t)(d + 1) << 32);
    st(&nb->home, home);
    __builtin_amdgcn_s_waitcnt(0);  // the new bucket is whole before it is published
    st(reinterpret_cast<uint64_t*>(&directory[home]), reinterpret_cast<uint64_t>(nb));
    // level only grows: a cached value at or above d + 1 needs no update
    if (ld_cached(&level) < d + 1)
      gallatin::utils::atomic_max(&level, (uint64_t)(d + 1));
    __builtin_amdgcn_s_waitcnt(0);
    st(&b->ctl, ((uint64_t)(d + 1) << 32) | 1);  // still locked
    __builtin_amdgcn_s_waitcnt(0);  // new depth visible before keys leave
#pragma unroll
    for (int s = 0; s < items_per_bucket; ++s)
      if ((moved >> s) & 1) st(&b->keys[s], (uint64_t)defaultKey);
    return true;
  }

  // extendible_ht.cuh:842-1115.  Insert or replace; false when the table
  // cannot hold the key.
  __device__ bool insert(Key key, Val val) {
    if ((uint64_t)key == (uint64_t)defaultKey || (uint64_t)key == (uint64_t)tombstoneKey)
      return false;
    const uint64_t h = hash(key);
    int state = 0;  // 0 pending, 1 stored, 2 refused
    for (uint32_t tries = 0; state == 0 && tries < kTries; ++tries) {
      bucket* b = owner(h, tries > 0);
      if (!b) {
        state = 2;
        break;
      }
      uint32_t d;
      if (!try_lock(b, d)) {
        __builtin_amdgcn_s_sleep(2);
        continue;
      }
      int match, free_slot;
      const uint64_t home = scan(b, (uint64_t)key, match, free_slot);
      if ((h & mask(d)) == home) {  // still the owner
        if (match >= 0) {
          st(&b->vals[match], (uint64_t)val);
          state = 1;
        } else if (free_slot >= 0) {
          st(&b->vals[free_slot], (uint64_t)val);
          __builtin_amdgcn_s_waitcnt(0);  // value before key
          st(&b->keys[free_slot], (uint64_t)key);
          state = 1;
        } else if (d < max_bits && split(b, d, home)) {  // full: split, try again
          d += 1;
        } else {
          state = 2;
        }
      }
      unlock(b, d);
    }
    if (state != 1) gallatin::utils::atomic_add(&failed, (uint64_t)1);
    return state == 1;
  }

  // extendible_ht.cuh:1291-1375.  The depth, home and keys are loaded
  // together; a missing key is reported only if the depth is unchanged
  // when looked at again after the keys (a split raises the depth before
  // it clears moved keys), and a found key's value is taken only if the
  // key is still in its slot after the value was read.
  __device__ bool query(Key key, Val& val) {
    const uint64_t h = hash(key);
    for (uint32_t tries = 0; tries < kTries; ++tries) {
      bucket* b = owner(h, tries > 0);
      if (!b) return false;
      const uint32_t d0 = depth_of(ld(&b->ctl));
      int match, empty;
      const uint64_t home = scan(b, (uint64_t)key, match, empty);
      if ((h & mask(d0)) != home) continue;  // being split: look again
      if (match >= 0) {
        const uint64_t v = ld(&b->vals[match]);
        __builtin_amdgcn_s_waitcnt(0);  // the value is read before the key is read again
        // Between the two reads the key may have left the slot (moved by a
        // split or removed) and another insert may have reused it, value
        // first: the value belongs to the key only if the key is still there.
        if (ld(&b->keys[match]) == (uint64_t)key) {
          val = (Val)v;
          return true;
        }
        continue;
      }
      if (depth_of(ld(&b->ctl)) == d0) return false;
    }
    return false;
  }

  // Remove key; true if it was present.  (No counterpart in the reference.)
  __device__ bool remove(Key key) {
    const uint64_t h = hash(key);
    int state = 0;  // 0 pending, 1 removed, 2 absent
    for (uint32_t tries = 0; state == 0 && tries < kTries; ++tries) {
      bucket* b = owner(h, tries > 0);
      if (!b) break;
      uint32_t d;
      if (!try_lock(b, d)) {
        __builtin_amdgcn_s_sleep(2);
        continue;
      }
      int match, empty;
      const uint64_t home = scan(b, (uint64_t)key, match, empty);
      if ((h & mask(d)) == home) {
        state = 2;
        if (match >= 0) {
          st(&b->keys[match], (uint64_t)defaultKey);
          state = 1;
        }
      }
      unlock(b, d);
    }
    return state == 1;
  }
};

}  // namespace data_structs
}  // namespace gallatin

#endif

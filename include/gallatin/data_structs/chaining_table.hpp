// Allocator-backed chained hash table (chaining_table.cuh:26-335).
//
// The table has nblocks = nslots / size chain heads; key k belongs to chain
// MurmurHash64A(&k, sizeof(Key), seed) % nblocks.  A chain is a list of
// blocks of `size` key/value pairs taken from the global allocator.
//
// insert walks the chain and claims the first empty slot with a
// compare-and-swap of its key, writes the value, then sets the slot's bit
// in the block's `ready` mask.  When every block of the chain is full it
// attaches a new block at the end: allocate, format, compare-and-swap the
// link (the loser returns its block).  Slots are never emptied, so a block
// is full before any insert moves past it, and a chain that received c keys
// has exactly ceil(c / size) blocks.  query walks the chain and returns the
// value of the first ready slot holding the key; the keys, the ready mask
// and the link of a block are loaded together.
//
// Same as the reference: duplicate keys are stored again and query returns
// the first in chain order; there is no delete; `defaultKey` marks an empty
// slot.  Differences: a value is published (ready bit) before queries see
// its key, where the reference writes the value after the key, so its
// concurrent queries can read an unwritten value; insert returns whether it
// stored the pair (false for defaultKey, or when no block could be
// allocated) instead of nothing; free_on_device also returns the chain
// blocks to the allocator (the reference frees only the head array); a
// table with nslots < size is refused (nullptr) instead of dividing by 0.
// Every block word is read and written with device-coherent atomics.
#ifndef GALLATIN_CHAINING_TABLE_HPP
#define GALLATIN_CHAINING_TABLE_HPP

#include <hip/hip_runtime.h>

#include <iostream>

#include "gallatin/allocators/global_allocator.hpp"
#include "gallatin/allocators/murmurhash.hpp"

namespace gallatin {
namespace data_structs {

// chaining_table.cuh:53-107.
template <typename Key, typename Val, int size>
struct chaining_block {
  static_assert(size >= 1 && size <= 64, "1 to 64 pairs per block");
  uint64_t next;   // address of the next block of the chain, 0 = none
  uint64_t ready;  // bit i: the value of slot i is written
  Key keys[size];
  Val vals[size];

  __device__ void init(Key defaultKey) {
    gallatin::utils::atomic_store(&next, (uint64_t)0);
    gallatin::utils::atomic_store(&ready, (uint64_t)0);
    for (int i = 0; i < size; ++i) gallatin::utils::atomic_store(&keys[i], defaultKey);
  }
};

template <typename table>
__global__ void chaining_table_fill_buffers(table* t) {
  const uint64_t i = gallatin::utils::get_tid();
  if (i < t->nblocks) t->attach_block(reinterpret_cast<uint64_t*>(&t->pointer_list[i]));
}

template <typename table>
__global__ void chaining_table_free_chains(table* t) {
  const uint64_t i = gallatin::utils::get_tid();
  if (i < t->nblocks) t->free_chain(i);
}

// Longest chain and total blocks, one atomic per wave for each.
template <typename table>
__global__ void calculate_chain_kernel(table* t, uint64_t* max_len, uint64_t* total,
                                       uint64_t nblocks) {
  const uint64_t i = gallatin::utils::get_tid();
  uint64_t len = i < nblocks ? t->chain_length(i) : 0;
  uint64_t mx = len;
  for (int off = 32; off >= 1; off >>= 1) {
    const int src = (int)(gallatin::wave::lane() ^ (uint32_t)off);
    len += gallatin::wave::fetch(len, src);
    const uint64_t o = gallatin::wave::fetch(mx, src);
    mx = o > mx ? o : mx;
  }
  if (gallatin::wave::lane() == 0) {
    if (len) gallatin::utils::atomic_add(total, len);
    if (mx) gallatin::utils::atomic_max(max_len, mx);
  }
}

template <typename table>
__global__ void chain_lengths_kernel(table* t, uint32_t* out, uint64_t nblocks) {
  const uint64_t i = gallatin::utils::get_tid();
  if (i < nblocks) out[i] = (uint32_t)t->chain_length(i);
}

template <typename Key, typename Val, int size = 4>
struct chaining_table {
  static_assert(sizeof(Key) == 4 || sizeof(Key) == 8, "32- or 64-bit keys");
  static_assert(sizeof(Val) == 4 || sizeof(Val) == 8, "32- or 64-bit values");
  using my_type = chaining_table<Key, Val, size>;
  using block_type = chaining_block<Key, Val, size>;
  using key_type = Key;
  using val_type = Val;
  static constexpr int block_pairs = size;
  static constexpr uint32_t kMaxHops = 1u << 24;  // bound on a chain walk

  uint64_t nslots;
  uint64_t nblocks;
  uint64_t seed;
  block_type** pointer_list;
  Key defaultKey;
  uint64_t failed;  // inserts refused for want of a block

  // ---- host ----
  // chaining_table.cuh:132-156.  Needs the global allocator.
  static __host__ my_type* generate_on_device(uint64_t ext_nslots, Key ext_defaultKey,
                                              uint64_t ext_seed) {
    if (ext_nslots / size == 0) return nullptr;
    my_type host{};
    host.seed = ext_seed;
    host.nslots = ext_nslots;
    host.nblocks = ext_nslots / size;
    host.defaultKey = ext_defaultKey;
    if (hipMalloc(&host.pointer_list, sizeof(block_type*) * host.nblocks) != hipSuccess)
      return nullptr;
    my_type* dev = nullptr;
    if (hipMemset(host.pointer_list, 0, sizeof(block_type*) * host.nblocks) != hipSuccess ||
        hipMalloc(&dev, sizeof(my_type)) != hipSuccess) {
      (void)hipFree(host.pointer_list);
      return nullptr;
    }
    if (hipMemcpy(dev, &host, sizeof(my_type), hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(host.pointer_list);
      (void)hipFree(dev);
      return nullptr;
    }
    return dev;
  }

  // chaining_table.cuh:158-187: every chain starts with one block.
  static __host__ my_type* generate_on_device_prealloc(uint64_t ext_nslots, Key ext_defaultKey,
                                                       uint64_t ext_seed) {
    my_type* dev = generate_on_device(ext_nslots, ext_defaultKey, ext_seed);
    if (!dev) return nullptr;
    const uint64_t nb = ext_nslots / size;
    chaining_table_fill_buffers<my_type><<<(unsigned)((nb + 255) / 256), 256>>>(dev);
    if (hipDeviceSynchronize() != hipSuccess) return nullptr;
    return dev;
  }

  // chaining_table.cuh:189-197, and the chain blocks.
  static __host__ void free_on_device(my_type* dev) {
    if (!dev) return;
    my_type host;
    if (hipMemcpy(&host, dev, sizeof(my_type), hipMemcpyDeviceToHost) != hipSuccess) return;
    chaining_table_free_chains<my_type><<<(unsigned)((host.nblocks + 255) / 256), 256>>>(dev);
    (void)hipDeviceSynchronize();
    (void)hipFree(host.pointer_list);
    (void)hipFree(dev);
  }

  // Longest chain and total blocks, in blocks; called on the device
  // pointer, as the reference's print_chain_stats is.
  __host__ void chain_stats(uint64_t* max_len, uint64_t* total) {
    my_type host;
    uint64_t* d = nullptr;
    uint64_t h[2] = {0, 0};
    if (hipMemcpy(&host, this, sizeof(my_type), hipMemcpyDeviceToHost) == hipSuccess &&
        hipMalloc(&d, sizeof(h)) == hipSuccess) {
      (void)hipMemset(d, 0, sizeof(h));
      calculate_chain_kernel<my_type><<<(unsigned)((host.nblocks + 255) / 256), 256>>>(
          this, d, d + 1, host.nblocks);
      (void)hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
      (void)hipFree(d);
    }
    if (max_len) *max_len = h[0];
    if (total) *total = h[1];
  }

  // chaining_table.cuh:295-322.
  __host__ void print_chain_stats() {
    my_type host;
    if (hipMemcpy(&host, this, sizeof(my_type), hipMemcpyDeviceToHost) != hipSuccess) return;
    uint64_t mx = 0, total = 0;
    chain_stats(&mx, &total);
    std::cout << "Chains - Max: " << mx << ", Avg: " << 1.0 * total / host.nblocks
              << ", nblocks: " << host.nblocks << std::endl;
  }

  // ---- device ----
  static __device__ __forceinline__ uint64_t ld64(const uint64_t* p) {
    return gallatin::utils::atomic_load(p);
  }

  __device__ uint64_t chain_of(Key k) const {
    return gallatin::hashers::MurmurHash64A(&k, (int)sizeof(Key), seed) % nblocks;
  }

  // chaining_table.cuh:202-219.  False when no block could be allocated
  // and the link is still empty.
  __device__ bool attach_block(uint64_t* link) {
    block_type* nb =
        static_cast<block_type*>(gallatin::allocators::global_malloc(sizeof(block_type)));
    if (!nb) return ld64(link) != 0;
    nb->init(defaultKey);
    __builtin_amdgcn_s_waitcnt(0);  // the block is formatted before it is linked
    if (gallatin::utils::atomic_cas(link, (uint64_t)0, reinterpret_cast<uint64_t>(nb)) != 0)
      gallatin::allocators::global_free(nb);
    return true;
  }

  // chaining_table.cuh:240-282.
  __device__ bool insert(Key newKey, Val newVal) {
    if (newKey == defaultKey) return false;
    uint64_t* link = reinterpret_cast<uint64_t*>(&pointer_list[chain_of(newKey)]);
    for (uint32_t hop = 0; hop < kMaxHops; ++hop) {
      uint64_t b = ld64(link);
      if (!b) {
        if (!attach_block(link)) break;
        b = ld64(link);
      }
      block_type* blk = reinterpret_cast<block_type*>(b);
      Key k[size];
#pragma unroll
      for (int i = 0; i < size; ++i) k[i] = gallatin::utils::atomic_load(&blk->keys[i]);
      for (int i = 0; i < size; ++i) {
        if (k[i] != defaultKey) continue;
        if (gallatin::utils::atomic_cas(&blk->keys[i], defaultKey, newKey) == defaultKey) {
          gallatin::utils::atomic_store(&blk->vals[i], newVal);
          __builtin_amdgcn_s_waitcnt(0);  // the value lands before the slot counts
          gallatin::utils::atomic_or(&blk->ready, (uint64_t)1 << i);
          return true;
        }
      }
      link = &blk->next;  // every slot was taken: the next block
    }
    gallatin::utils::atomic_add(&failed, (uint64_t)1);
    return false;
  }

  // chaining_table.cuh:284-309.
  __device__ bool query(Key queryKey, Val& returnVal) {
    if (queryKey == defaultKey) return false;
    uint64_t b = ld64(reinterpret_cast<uint64_t*>(&pointer_list[chain_of(queryKey)]));
    for (uint32_t hop = 0; b && hop < kMaxHops; ++hop) {
      block_type* blk = reinterpret_cast<block_type*>(b);
      Key k[size];
#pragma unroll
      for (int i = 0; i < size; ++i) k[i] = gallatin::utils::atomic_load(&blk->keys[i]);
      const uint64_t rdy = ld64(&blk->ready);
      const uint64_t nxt = ld64(&blk->next);
      for (int i = 0; i < size; ++i) {
        if (k[i] == queryKey && ((rdy >> i) & 1)) {
          returnVal = gallatin::utils::atomic_load(&blk->vals[i]);
          return true;
        }
      }
      b = nxt;
    }
    return false;
  }

  // chaining_table.cuh:220-236: blocks in chain i.
  __device__ uint64_t chain_length(uint64_t i) {
    uint64_t n = 0;
    uint64_t b = ld64(reinterpret_cast<uint64_t*>(&pointer_list[i]));
    while (b && n < kMaxHops) {
      ++n;
      b = ld64(&reinterpret_cast<block_type*>(b)->next);
    }
    return n;
  }

  // Returns the blocks of chain i to the allocator (no concurrent users).
  __device__ void free_chain(uint64_t i) {
    uint64_t b = reinterpret_cast<uint64_t>(pointer_list[i]);
    pointer_list[i] = nullptr;
    for (uint32_t hop = 0; b && hop < kMaxHops; ++hop) {
      block_type* blk = reinterpret_cast<block_type*>(b);
      b = blk->next;
      gallatin::allocators::global_free(blk);
    }
  }
};

}  // namespace data_structs
}  // namespace gallatin

#endif

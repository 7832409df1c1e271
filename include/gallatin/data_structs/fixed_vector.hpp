// Allocator-backed growable vector (fixed_vector.cuh:72-512).
//
// Items live in a directory of chunks taken from the global allocator:
// chunk 0 holds min_items items, chunk i >= 1 holds min_items << (i - 1)
// (positions [min_items << (i - 1), min_items << i)), up to max_items in
// all.  insert() takes the next position with one atomic and allocates a
// chunk the first time a position in it is taken; other inserters into the
// same chunk wait for its pointer (a bounded wait here, where the reference
// spins forever when the allocation fails).  bulk_insert() takes the
// positions of all active lanes of the wave with one atomic, in lane order
// (the reference's coalesced-group form).
//
// Same as the reference: not safe to read an item while another thread may
// still be inserting it (readers run in a later kernel).  Differences: an
// insert that finds max_items reached, or whose chunk cannot be allocated,
// returns ~0; free_device_vector also releases the vector object; host
// (on_host = true) backing is not supported.
#ifndef GALLATIN_FIXED_VECTOR_HPP
#define GALLATIN_FIXED_VECTOR_HPP

#include <hip/hip_runtime.h>

#include <vector>

#include "gallatin/allocators/global_allocator.hpp"

namespace gallatin {
namespace data_structs {

// alloc_utils.cuh:614-617: bits needed to represent x.
constexpr uint64_t number_of_bits(uint64_t x) { return x < 2 ? x : 1 + number_of_bits(x >> 1); }

template <typename vector>
__global__ void init_dev_vector(vector* vec) {
  if (blockIdx.x == 0 && threadIdx.x == 0) vec->init();
}

template <typename vector>
__global__ void free_dev_vector(vector* vec) {
  if (blockIdx.x == 0 && threadIdx.x == 0) vec->free_vector();
}

template <typename T, typename vector>
__global__ void copy_vector_to_dev_array(T* dev_array, vector* dev_vector, uint64_t size) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid < size) dev_array[tid] = (*dev_vector)[tid];
}

template <typename T, uint64_t min_items, uint64_t max_items, bool on_host = false>
struct fixed_vector {
  using my_type = fixed_vector<T, min_items, max_items, on_host>;
  static_assert(!on_host, "host-memory backing is not supported");
  static_assert(min_items > 0 && (min_items & (min_items - 1)) == 0, "min_items: power of two");
  static_assert(max_items >= min_items && (max_items & (max_items - 1)) == 0,
                "max_items: power of two, at least min_items");

  static constexpr uint32_t n_directory =
      (uint32_t)(number_of_bits(max_items - 1) + 1 - number_of_bits(min_items - 1));
  static constexpr uint32_t kWaitCap = 1u << 22;

  T* directory[n_directory];
  uint64_t locks;  // bit i: chunk i is being (or has been) allocated
  uint64_t size;   // positions handed out (may exceed max_items)

  // ---- host ----
  static __host__ my_type* get_device_vector() {
    my_type* dev = nullptr;
    if (hipMalloc(&dev, sizeof(my_type)) != hipSuccess) return nullptr;
    init_dev_vector<my_type><<<1, 1>>>(dev);
    if (hipDeviceSynchronize() != hipSuccess) return nullptr;
    return dev;
  }

  static __host__ void free_device_vector(my_type* dev) {
    if (!dev) return;
    free_dev_vector<my_type><<<1, 1>>>(dev);
    (void)hipDeviceSynchronize();
    (void)hipFree(dev);
  }

  // Items [0, size) copied to the host (fixed_vector.cuh:475-506).
  static __host__ std::vector<T> export_to_host(my_type* dev) {
    uint64_t n = 0;
    std::vector<T> out;
    if (hipMemcpy(&n, &dev->size, sizeof(n), hipMemcpyDeviceToHost) != hipSuccess) return out;
    if (n > max_items) n = max_items;
    out.resize(n);
    if (n == 0) return out;
    T* tmp = nullptr;
    if (hipMalloc(&tmp, sizeof(T) * n) != hipSuccess) return std::vector<T>();
    copy_vector_to_dev_array<T, my_type><<<(unsigned)((n + 255) / 256), 256>>>(tmp, dev, n);
    if (hipMemcpy(out.data(), tmp, sizeof(T) * n, hipMemcpyDeviceToHost) != hipSuccess)
      out.clear();
    (void)hipFree(tmp);
    return out;
  }

  // ---- device ----
  static __device__ __host__ constexpr uint64_t chunk_items(uint32_t i) {
    return i <= 1 ? min_items : min_items << (i - 1);
  }

  // Chunk of position p (fixed_vector.cuh:164-186).
  static __device__ __host__ uint32_t get_directory_index(uint64_t p) {
    if (p < min_items) return 0;
    return (uint32_t)(number_of_bits(p / min_items));
  }

  // Position inside its chunk (fixed_vector.cuh:189-200).
  static __device__ __host__ uint64_t get_local_position(uint64_t p, uint32_t i) {
    return i == 0 ? p : p - (min_items << (i - 1));
  }

  __device__ void init() {
    size = 0;
    locks = 0;
    for (uint32_t i = 0; i < n_directory; ++i) directory[i] = nullptr;
    add_new_backing(0);
  }

  // fixed_vector.cuh:92-110: the first thread to claim chunk i allocates it.
  // 1: this call allocated the chunk; 0: another thread holds (or held) the
  // claim; -1: this call held the claim and the allocation failed, so the
  // claim is given up again (a later call may retry).
  __device__ int add_new_backing(uint32_t i) {
    const uint64_t bit = 1ULL << i;
    if (gallatin::utils::atomic_or(&locks, bit) & bit) return 0;
    T* b = static_cast<T*>(gallatin::allocators::global_malloc(sizeof(T) * chunk_items(i)));
    if (!b) {
      gallatin::utils::atomic_and(&locks, ~bit);
      return -1;
    }
    gallatin::utils::atomic_store(reinterpret_cast<uint64_t*>(&directory[i]),
                                  reinterpret_cast<uint64_t>(b));
    return 1;
  }

  // The chunk's memory, or nullptr when it cannot be had: this call's own
  // allocation failed, or the thread that claimed the chunk gave the claim
  // up without a pointer (its allocation failed), or the bounded wait ran out.
  __device__ T* chunk(uint32_t i) {
    uint64_t* slot = reinterpret_cast<uint64_t*>(&directory[i]);
    uint64_t b = gallatin::utils::atomic_load(slot);
    // The lanes of this wave that find the same chunk (same vector, same
    // index) missing let their lowest lane try the claim for all of them,
    // one chunk after another, before any of them polls.  A lane must never
    // poll for a claim held by another lane of its own wave: the wave runs
    // one side of the branch at a time, and if it ran the polling lanes
    // first they would wait out the whole poll while the claimer sat masked
    // off (measured: a full wave took 1 s, one lane 3 ms).
    const bool need = b == 0;
    const uint64_t key = reinterpret_cast<uint64_t>(slot);
    int mine = 0;
    for (uint64_t rem = __ballot(need); rem;) {
      const int l = gallatin::wave::first(rem);
      const bool same = need && key == gallatin::wave::readlane(key, l);
      const uint64_t grp = __ballot(same);
      uint32_t r = 0;
      if ((int)gallatin::wave::lane() == l) r = (uint32_t)add_new_backing(i);
      r = gallatin::wave::readlane(r, l);
      if (same) mine = (int)r;
      rem &= ~grp;
    }
    if (!need) return reinterpret_cast<T*>(b);
    if (mine < 0) return nullptr;
    const uint64_t bit = 1ULL << i;
    for (uint32_t w = 0; w < kWaitCap; ++w) {
      b = gallatin::utils::atomic_load(slot);
      if (b) break;
      if (!(gallatin::utils::atomic_load(&locks) & bit)) break;  // claimer failed
      __builtin_amdgcn_s_sleep(2);
    }
    return reinterpret_cast<T*>(b);
  }

  __device__ uint64_t place(uint64_t p, const T& item) {
    if (p >= max_items) return ~0ULL;
    const uint32_t d = get_directory_index(p);
    T* c = chunk(d);
    if (!c) return ~0ULL;
    c[get_local_position(p, d)] = item;
    return p;
  }

  // fixed_vector.cuh:216-257.  The position written, or ~0.
  __device__ uint64_t insert(T item) {
    return place(gallatin::utils::atomic_add(&size, (uint64_t)1), item);
  }

  // fixed_vector.cuh:261-310: one atomic for all active lanes of the wave.
  __device__ uint64_t bulk_insert(T item) {
    const uint64_t act = gallatin::wave::active();
    const int leader = gallatin::wave::first(act);
    uint64_t base = 0;
    if ((int)gallatin::wave::lane() == leader)
      base = gallatin::utils::atomic_add(&size, (uint64_t)gallatin::wave::count(act));
    base = gallatin::wave::readfirst(base);
    return place(base + gallatin::wave::count(act & gallatin::wave::lanes_below()), item);
  }

  __device__ T& operator[](uint64_t p) {
    const uint32_t d = get_directory_index(p);
    return directory[d][get_local_position(p, d)];
  }

  // fixed_vector.cuh:148-159: release every chunk.
  __device__ void free_vector() {
    for (uint32_t i = 0; i < n_directory; ++i) {
      if (directory[i]) gallatin::allocators::global_free(directory[i]);
      directory[i] = nullptr;
    }
    locks = 0;
    size = 0;
  }
};

}  // namespace data_structs
}  // namespace gallatin

#endif

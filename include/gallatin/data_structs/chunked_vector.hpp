// Allocator-backed vector whose buffer grows in power-of-two chunks up to a
// largest chunk, then in chunks of that largest size.
//
// The first chunks are fixed_vector's directory (fixed_vector.cuh:72-205):
// chunk 0 holds min_items items, chunk d >= 1 holds min_items << (d - 1), up
// to chunk n_grow - 1 of max_chunk_items.  Past those grow_items positions
// the vector keeps growing in chunks of max_chunk_items, whose pointers live
// in one index block taken from the allocator the first time a tail chunk
// is needed (tail_chunks pointers).  A session buffer of 16-item (128 B)
// first chunks and 8192-item (64 KiB) largest chunks therefore grows without
// a cap short of 16384 + tail_chunks * 8192 items, where fixed_vector stops
// at max_items.  No reference counterpart beyond fixed_vector's doubling
// chunks.
//
// One writer at a time: a vector is grown and written by one wave (or one
// thread), as a serving session is; readers run after the writer's kernel.
// Chunks are allocated by the caller (any allocator call of its choice, so a
// wave can allocate the next chunks of many vectors in one call) and
// installed with set_chunk; free_all returns every chunk and the index with
// global_free.
#ifndef GALLATIN_CHUNKED_VECTOR_HPP
#define GALLATIN_CHUNKED_VECTOR_HPP

#include <hip/hip_runtime.h>

#include "gallatin/allocators/global_allocator.hpp"
#include "gallatin/data_structs/fixed_vector.hpp"

namespace gallatin {
namespace data_structs {

template <typename T, uint64_t min_items, uint64_t max_chunk_items, uint32_t tail_chunks>
struct chunked_vector {
  static_assert(min_items > 0 && (min_items & (min_items - 1)) == 0, "min_items: power of two");
  static_assert(max_chunk_items >= min_items && (max_chunk_items & (max_chunk_items - 1)) == 0,
                "max_chunk_items: power of two, at least min_items");
  static_assert(tail_chunks > 0, "at least one tail chunk");

  // doubling chunks 0 .. n_grow - 1 (the last one holds max_chunk_items)
  static constexpr uint32_t n_grow = (uint32_t)number_of_bits(max_chunk_items / min_items) + 1;
  static constexpr uint64_t grow_items = min_items << (n_grow - 1);
  static constexpr uint32_t n_chunks = n_grow + tail_chunks;
  static constexpr uint64_t max_items = grow_items + (uint64_t)tail_chunks * max_chunk_items;
  static constexpr uint64_t index_bytes = sizeof(T*) * tail_chunks;

  T* grow[n_grow];
  T** tail;       // tail_chunks pointers, or nullptr before the first tail chunk
  uint64_t size;  // items written

  // ---- chunk geometry (host and device) ----
  static __host__ __device__ constexpr uint64_t chunk_items(uint32_t d) {
    return d <= 1 ? min_items : (d < n_grow ? min_items << (d - 1) : max_chunk_items);
  }
  static __host__ __device__ constexpr uint64_t chunk_start(uint32_t d) {
    return d == 0 ? 0 : (d < n_grow ? min_items << (d - 1)
                                    : grow_items + (uint64_t)(d - n_grow) * max_chunk_items);
  }
  static __host__ __device__ uint32_t chunk_of(uint64_t p) {
    if (p >= grow_items) return n_grow + (uint32_t)((p - grow_items) / max_chunk_items);
    if (p < min_items) return 0;
    return (uint32_t)number_of_bits(p / min_items);
  }

  // ---- device, single writer ----
  __device__ void init() {
    for (uint32_t d = 0; d < n_grow; ++d) grow[d] = nullptr;
    tail = nullptr;
    size = 0;
  }

  // Chunk d's memory, or nullptr when it has none yet.
  __device__ T* chunk(uint32_t d) const {
    if (d < n_grow) return grow[d];
    return tail ? tail[d - n_grow] : nullptr;
  }

  // Tail chunks need the index first (needs_index, then set_index).
  __device__ bool needs_index(uint32_t d) const { return d >= n_grow && tail == nullptr; }
  __device__ void set_index(T** idx) {
    for (uint32_t i = 0; i < tail_chunks; ++i) idx[i] = nullptr;
    tail = idx;
  }
  __device__ void set_chunk(uint32_t d, T* c) {
    if (d < n_grow)
      grow[d] = c;
    else
      tail[d - n_grow] = c;
  }

  __device__ T& operator[](uint64_t p) {
    const uint32_t d = chunk_of(p);
    return chunk(d)[p - chunk_start(d)];
  }

  // Chunks held (the index not counted).
  __device__ uint32_t chunks_held() const {
    uint32_t n = 0;
    for (uint32_t d = 0; d < n_grow; ++d) n += grow[d] != nullptr;
    if (tail)
      for (uint32_t i = 0; i < tail_chunks; ++i) n += tail[i] != nullptr;
    return n;
  }

  // Every chunk and the index back to the global allocator.  Lanes of a
  // wave may free their own vectors together: the frees of chunk d of all
  // of them are one global_free call.
  __device__ void free_all() {
    for (uint32_t d = 0; d < n_grow; ++d) {
      if (grow[d]) gallatin::allocators::global_free(grow[d]);
      grow[d] = nullptr;
    }
    if (tail) {
      // tail chunks are installed in order, so the first empty entry ends them
      for (uint32_t i = 0; i < tail_chunks && tail[i]; ++i)
        gallatin::allocators::global_free(tail[i]);
      gallatin::allocators::global_free(tail);
      tail = nullptr;
    }
    size = 0;
  }

  // free_all by a whole wave: every active lane calls it for the same
  // vector, and lane i frees chunks i, i + 64, ... in one global_free call
  // of the wave per 64 chunks, where free_all makes one call per chunk.
  // Returns the chunks freed (the index not counted), on every lane.
  __device__ __noinline__ uint32_t free_all_wave() {
    static_assert(n_grow <= 64, "one lane per doubling chunk");
    const uint32_t lane = __lane_id();
    T* g = lane < n_grow ? grow[lane] : nullptr;
    uint32_t n = (uint32_t)__builtin_popcountll(__ballot(g != nullptr));
    if (g) gallatin::allocators::global_free(g);
    T** t = tail;
    if (t) {
      // tail chunks are installed in order, so the first empty entry ends them
      for (uint32_t base = 0; base < tail_chunks; base += 64) {
        const bool in = base + lane < tail_chunks;
        T* c = in ? t[base + lane] : nullptr;
        n += (uint32_t)__builtin_popcountll(__ballot(c != nullptr));
        if (c) gallatin::allocators::global_free(c);
        if (__ballot(in && c == nullptr)) break;
      }
      if (lane == (uint32_t)__builtin_ctzll(__ballot(1))) gallatin::allocators::global_free(t);
    }
    if (lane < n_grow) grow[lane] = nullptr;
    if (lane == 0) {
      tail = nullptr;
      size = 0;
    }
    return n;
  }
};

}  // namespace data_structs
}  // namespace gallatin

#endif

// The library's global allocator instance and its bulk kernels
// (global_allocator.cuh:41-97), and the allocator-backed data structures
// that allocate through it (fixed_vector, extendible_hash_table,
// chaining_table, block_queue).  This is
// the only translation unit of the library that includes
// global_allocator.hpp, which defines the device symbol global_gallatin.
#include <hip/hip_runtime.h>

#include <numeric>

#include <type_traits>

#include "gallatin/allocators/global_allocator.hpp"
#include "gallatin/data_structs/block_queue.hpp"
#include "gallatin/data_structs/chaining_table.hpp"
#include "gallatin/data_structs/chunked_vector.hpp"
#include "gallatin/data_structs/extendible_ht.hpp"
#include "gallatin/data_structs/fixed_vector.hpp"
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

__global__ void k_global_malloc_fused_bulk(const uint64_t* sizes, uint64_t n, void** out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = global_malloc_fused(sizes[i]);
}

__global__ void k_global_free_fused_bulk(void* const* ptrs, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && ptrs[i]) global_free_fused(ptrs[i]);
}

__global__ void k_global_malloc_combined_bulk(const uint64_t* sizes, uint64_t n, int on_host,
                                              void** out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = global_malloc_combined(sizes[i], on_host != 0);
}

__global__ void k_global_free_combined_bulk(void* const* ptrs, uint64_t n, int on_host) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && ptrs[i]) global_free_combined(ptrs[i], on_host != 0);
}

__global__ void k_global_malloc_poison_bulk(const uint64_t* sizes, uint64_t n, void** out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = global_malloc_poison(sizes[i]);
}

__global__ void k_global_check_poison_bulk(void* const* ptrs, uint64_t n, uint8_t* ok) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ok[i] = global_check_poison(ptrs[i]) ? 1 : 0;
}

__global__ void k_global_free_poison_bulk(void* const* ptrs, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && ptrs[i]) global_free_poison(ptrs[i]);
}

static global_allocator_type* current_global_host() {
  global_allocator_type* p = nullptr;
  if (hipMemcpyFromSymbol(&p, HIP_SYMBOL(global_host_gallatin), sizeof(p)) != hipSuccess)
    return nullptr;
  return p;
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

int gal_init_global_host(uint64_t num_bytes, uint64_t seed, int print_info) {
  if (current_global_host()) return fail(GAL_ERR_INVALID, "host allocator already initialised");
  init_global_allocator_host(num_bytes, seed, print_info != 0, false);
  if (!current_global_host()) return fail(GAL_ERR_BOOT, "init_global_allocator_host failed");
  return GAL_OK;
}

int gal_free_global_host(void) {
  if (!current_global_host()) return fail(GAL_ERR_INVALID, "no host allocator");
  free_global_allocator_host();
  return GAL_OK;
}

int gal_print_global_stats_host(void) {
  if (!current_global_host()) return fail(GAL_ERR_INVALID, "no host allocator");
  print_global_stats_host();
  return GAL_OK;
}

gal_allocator* gal_global_host_handle(void) {
  return gal_from_device_state(current_global_host());
}

static unsigned grid_of(uint64_t n) { return (unsigned)((n + 255) / 256); }

int gal_global_malloc_fused_bulk(const uint64_t* sizes, uint64_t n, void** out, void* stream) {
  if ((!sizes || !out) && n) return fail(GAL_ERR_INVALID, "null argument");
  if (!current_global() || !current_global_host())
    return fail(GAL_ERR_INVALID, "fused malloc needs the global and the host allocator");
  if (n == 0) return GAL_OK;
  k_global_malloc_fused_bulk<<<grid_of(n), 256, 0, (hipStream_t)stream>>>(sizes, n, out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GAL_OK : hip_fail(e, "global_malloc_fused");
}

int gal_global_free_fused_bulk(void* const* ptrs, uint64_t n, void* stream) {
  if (!ptrs && n) return fail(GAL_ERR_INVALID, "null argument");
  if (!current_global() || !current_global_host())
    return fail(GAL_ERR_INVALID, "fused free needs the global and the host allocator");
  if (n == 0) return GAL_OK;
  k_global_free_fused_bulk<<<grid_of(n), 256, 0, (hipStream_t)stream>>>(ptrs, n);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GAL_OK : hip_fail(e, "global_free_fused");
}

int gal_global_malloc_combined_bulk(const uint64_t* sizes, uint64_t n, int on_host, void** out,
                                    void* stream) {
  if ((!sizes || !out) && n) return fail(GAL_ERR_INVALID, "null argument");
  if (!(on_host ? current_global_host() : current_global()))
    return fail(GAL_ERR_INVALID, "that allocator is not initialised");
  if (n == 0) return GAL_OK;
  k_global_malloc_combined_bulk<<<grid_of(n), 256, 0, (hipStream_t)stream>>>(sizes, n, on_host,
                                                                            out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GAL_OK : hip_fail(e, "global_malloc_combined");
}

int gal_global_free_combined_bulk(void* const* ptrs, uint64_t n, int on_host, void* stream) {
  if (!ptrs && n) return fail(GAL_ERR_INVALID, "null argument");
  if (!(on_host ? current_global_host() : current_global()))
    return fail(GAL_ERR_INVALID, "that allocator is not initialised");
  if (n == 0) return GAL_OK;
  k_global_free_combined_bulk<<<grid_of(n), 256, 0, (hipStream_t)stream>>>(ptrs, n, on_host);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GAL_OK : hip_fail(e, "global_free_combined");
}

int gal_global_malloc_poison_bulk(const uint64_t* sizes, uint64_t n, void** out, void* stream) {
  if ((!sizes || !out) && n) return fail(GAL_ERR_INVALID, "null argument");
  if (!current_global()) return fail(GAL_ERR_INVALID, "no global allocator");
  if (n == 0) return GAL_OK;
  k_global_malloc_poison_bulk<<<grid_of(n), 256, 0, (hipStream_t)stream>>>(sizes, n, out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GAL_OK : hip_fail(e, "global_malloc_poison");
}

int gal_global_check_poison_bulk(void* const* ptrs, uint64_t n, uint8_t* ok, void* stream) {
  if ((!ptrs || !ok) && n) return fail(GAL_ERR_INVALID, "null argument");
  if (!current_global()) return fail(GAL_ERR_INVALID, "no global allocator");
  if (n == 0) return GAL_OK;
  k_global_check_poison_bulk<<<grid_of(n), 256, 0, (hipStream_t)stream>>>(ptrs, n, ok);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GAL_OK : hip_fail(e, "global_check_poison");
}

int gal_global_free_poison_bulk(void* const* ptrs, uint64_t n, void* stream) {
  if (!ptrs && n) return fail(GAL_ERR_INVALID, "null argument");
  if (!current_global()) return fail(GAL_ERR_INVALID, "no global allocator");
  if (n == 0) return GAL_OK;
  k_global_free_poison_bulk<<<grid_of(n), 256, 0, (hipStream_t)stream>>>(ptrs, n);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GAL_OK : hip_fail(e, "global_free_poison");
}

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

// ---- allocator-backed data structures ----

using gallatin::data_structs::extendible_hash_table;
using gallatin::data_structs::fixed_vector;

// vector of 64-bit items: 64 .. 2^22 items
using vec_u64 = fixed_vector<uint64_t, 64, 1ULL << 22>;
// hash tables of 64-bit keys / values, 15 pairs (256 B) per bucket:
// small (2^4 .. 2^20 buckets, splits early) and large (2^16 .. 2^28)
using table_small = extendible_hash_table<uint64_t, 0ULL, ~0ULL, uint64_t, 15, 4, 20>;
using table_large = extendible_hash_table<uint64_t, 0ULL, ~0ULL, uint64_t, 15, 16, 28>;

struct gal_vector {
  vec_u64* dev = nullptr;
};

struct gal_table {
  int kind = 0;
  void* dev = nullptr;
};

namespace {

__global__ void k_vec_insert(vec_u64* v, const uint64_t* items, uint64_t n, uint64_t* out,
                             int bulk) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t r = bulk ? v->bulk_insert(items[i]) : v->insert(items[i]);
  if (out) out[i] = r;
}

template <typename T>
__global__ void k_ht_insert(T* t, const uint64_t* keys, const uint64_t* vals, uint64_t n,
                            uint8_t* ok) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool r = t->insert(keys[i], vals[i]);
  if (ok) ok[i] = r;
}

template <typename T>
__global__ void k_ht_query(T* t, const uint64_t* keys, uint64_t n, uint64_t* vals,
                           uint8_t* found) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t v = 0;
  const bool r = t->query(keys[i], v);
  if (vals) vals[i] = r ? v : 0;
  if (found) found[i] = r;
}

template <typename T>
__global__ void k_ht_remove(T* t, const uint64_t* keys, uint64_t n, uint8_t* ok) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool r = t->remove(keys[i]);
  if (ok) ok[i] = r;
}

unsigned blocks_for(uint64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" {

int gal_vector_create(gal_vector** out) {
  if (!out) return fail(GAL_ERR_INVALID, "null argument");
  if (!current_global()) return fail(GAL_ERR_INVALID, "no global allocator");
  vec_u64* dev = vec_u64::get_device_vector();
  if (!dev) return fail(GAL_ERR_HIP, "vector creation failed");
  *out = new gal_vector{dev};
  return GAL_OK;
}

int gal_vector_destroy(gal_vector* v) {
  if (!v) return GAL_OK;
  vec_u64::free_device_vector(v->dev);
  delete v;
  return GAL_OK;
}

int gal_vector_insert(gal_vector* v, const uint64_t* items, uint64_t n, uint64_t* out_positions,
                      int bulk, void* stream) {
  if (!v || (!items && n)) return fail(GAL_ERR_INVALID, "null argument");
  if (n == 0) return GAL_OK;
  k_vec_insert<<<blocks_for(n), 256, 0, (hipStream_t)stream>>>(v->dev, items, n, out_positions,
                                                                 bulk);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GAL_OK : hip_fail(e, "vector insert");
}

int64_t gal_vector_export(gal_vector* v, uint64_t* out, uint64_t cap) {
  if (!v) return fail(GAL_ERR_INVALID, "null argument");
  if (hipDeviceSynchronize() != hipSuccess) return fail(GAL_ERR_HIP, "sync");
  std::vector<uint64_t> items = vec_u64::export_to_host(v->dev);
  if (out) {
    const uint64_t n = items.size() < cap ? items.size() : cap;
    for (uint64_t i = 0; i < n; ++i) out[i] = items[i];
  }
  return (int64_t)items.size();
}

int gal_table_create(int kind, uint64_t seed, gal_table** out) {
  if (!out || (kind != 0 && kind != 1)) return fail(GAL_ERR_INVALID, "bad table arguments");
  if (!current_global()) return fail(GAL_ERR_INVALID, "no global allocator");
  void* dev = kind == 0 ? (void*)table_small::generate_on_device(seed)
                        : (void*)table_large::generate_on_device(seed);
  if (!dev) return fail(GAL_ERR_HIP, "table creation failed (pool too small?)");
  *out = new gal_table{kind, dev};
  return GAL_OK;
}

int gal_table_destroy(gal_table* t) {
  if (!t) return GAL_OK;
  if (t->kind == 0)
    table_small::free_on_device(static_cast<table_small*>(t->dev));
  else
    table_large::free_on_device(static_cast<table_large*>(t->dev));
  delete t;
  return GAL_OK;
}

#define GAL_TABLE_DISPATCH(t, KERNEL, ...)                                                   \
  do {                                                                                        \
    if ((t)->kind == 0)                                                                       \
      KERNEL<table_small><<<blocks_for(n), 256, 0, (hipStream_t)stream>>>(                    \
          static_cast<table_small*>((t)->dev), __VA_ARGS__);                                  \
    else                                                                                      \
      KERNEL<table_large><<<blocks_for(n), 256, 0, (hipStream_t)stream>>>(                    \
          static_cast<table_large*>((t)->dev), __VA_ARGS__);                                  \
  } while (0)

int gal_table_insert(gal_table* t, const uint64_t* keys, const uint64_t* vals, uint64_t n,
                     uint8_t* ok, void* stream) {
  if (!t || ((!keys || !vals) && n)) return fail(GAL_ERR_INVALID, "null argument");
  if (n == 0) return GAL_OK;
  GAL_TABLE_DISPATCH(t, k_ht_insert, keys, vals, n, ok);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GAL_OK : hip_fail(e, "table insert");
}

int gal_table_query(gal_table* t, const uint64_t* keys, uint64_t n, uint64_t* vals,
                    uint8_t* found, void* stream) {
  if (!t || (!keys && n)) return fail(GAL_ERR_INVALID, "null argument");
  if (n == 0) return GAL_OK;
  GAL_TABLE_DISPATCH(t, k_ht_query, keys, n, vals, found);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GAL_OK : hip_fail(e, "table query");
}

int gal_table_remove(gal_table* t, const uint64_t* keys, uint64_t n, uint8_t* ok,
                     void* stream) {
  if (!t || (!keys && n)) return fail(GAL_ERR_INVALID, "null argument");
  if (n == 0) return GAL_OK;
  GAL_TABLE_DISPATCH(t, k_ht_remove, keys, n, ok);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GAL_OK : hip_fail(e, "table remove");
}

#undef GAL_TABLE_DISPATCH

int gal_table_stats(gal_table* t, uint64_t* keys, uint64_t* buckets, uint64_t* level,
                    uint64_t* failed) {
  if (!t) return fail(GAL_ERR_INVALID, "null argument");
  if (hipDeviceSynchronize() != hipSuccess) return fail(GAL_ERR_HIP, "sync");
  uint64_t hdr[5] = {0, 0, 0, 0, 0};  // directory, level, seed, reserved, failed
  if (hipMemcpy(hdr, t->dev, sizeof(hdr), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(GAL_ERR_HIP, "stats copy");
  uint64_t nb = 0;
  const uint64_t nk = t->kind == 0
                          ? table_small::count_keys(static_cast<table_small*>(t->dev), &nb)
                          : table_large::count_keys(static_cast<table_large*>(t->dev), &nb);
  if (keys) *keys = nk;
  if (buckets) *buckets = nb;
  if (level) *level = hdr[1];
  if (failed) *failed = hdr[4];
  return GAL_OK;
}

}  // extern "C"

// ---- chained hash tables ----
//
// Instantiations reachable through the C ABI, by kind: 64-bit keys and
// values with 4 (the reference's default), 16 and 1 pairs per block, and
// 32-bit keys and values with 4 and 32 pairs per block.

using gallatin::data_structs::chaining_table;
using chain_u64_4 = chaining_table<uint64_t, uint64_t, 4>;
using chain_u64_16 = chaining_table<uint64_t, uint64_t, 16>;
using chain_u32_4 = chaining_table<uint32_t, uint32_t, 4>;
using chain_u32_32 = chaining_table<uint32_t, uint32_t, 32>;
using chain_u64_1 = chaining_table<uint64_t, uint64_t, 1>;

struct gal_chain {
  int kind = 0;
  void* dev = nullptr;
  uint64_t nblocks = 0;
};

namespace {

// Calls f with a null pointer of the kind's table type.
template <typename F>
int with_chain_kind(int kind, F&& f) {
  switch (kind) {
    case 0: return f(static_cast<chain_u64_4*>(nullptr));
    case 1: return f(static_cast<chain_u64_16*>(nullptr));
    case 2: return f(static_cast<chain_u32_4*>(nullptr));
    case 3: return f(static_cast<chain_u32_32*>(nullptr));
    case 4: return f(static_cast<chain_u64_1*>(nullptr));
    default: return fail(GAL_ERR_INVALID, "unknown chaining table kind");
  }
}

template <typename T>
__global__ void k_chain_insert(T* t, const typename T::key_type* keys,
                               const typename T::val_type* vals, uint64_t n, uint8_t* ok) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool r = t->insert(keys[i], vals[i]);
  if (ok) ok[i] = r;
}

template <typename T>
__global__ void k_chain_query(T* t, const typename T::key_type* keys, uint64_t n,
                              typename T::val_type* vals, uint8_t* found) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  typename T::val_type v = 0;
  const bool r = t->query(keys[i], v);
  if (vals) vals[i] = r ? v : 0;
  if (found) found[i] = r;
}

}  // namespace

extern "C" {

int gal_chain_create(int kind, uint64_t nslots, uint64_t default_key, uint64_t seed,
                     int prealloc, gal_chain** out) {
  if (!out) return fail(GAL_ERR_INVALID, "null argument");
  if (!current_global()) return fail(GAL_ERR_INVALID, "no global allocator");
  return with_chain_kind(kind, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    using K = typename T::key_type;
    if (sizeof(K) == 4 && default_key > 0xFFFFFFFFULL)
      return fail(GAL_ERR_INVALID, "default key does not fit a 32-bit key");
    T* dev = prealloc ? T::generate_on_device_prealloc(nslots, (K)default_key, seed)
                      : T::generate_on_device(nslots, (K)default_key, seed);
    if (!dev) return fail(GAL_ERR_INVALID, "chaining table creation failed (nslots < block size?)");
    *out = new gal_chain{kind, dev, nslots / T::block_pairs};
    return GAL_OK;
  });
}

int gal_chain_destroy(gal_chain* c) {
  if (!c) return GAL_OK;
  const int r = with_chain_kind(c->kind, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    T::free_on_device(static_cast<T*>(c->dev));
    return GAL_OK;
  });
  delete c;
  return r;
}

int gal_chain_insert(gal_chain* c, const void* keys, const void* vals, uint64_t n, uint8_t* ok,
                     void* stream) {
  if (!c || ((!keys || !vals) && n)) return fail(GAL_ERR_INVALID, "null argument");
  if (n == 0) return GAL_OK;
  return with_chain_kind(c->kind, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    k_chain_insert<T><<<blocks_for(n), 256, 0, (hipStream_t)stream>>>(
        static_cast<T*>(c->dev), static_cast<const typename T::key_type*>(keys),
        static_cast<const typename T::val_type*>(vals), n, ok);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GAL_OK : hip_fail(e, "chain insert");
  });
}

int gal_chain_query(gal_chain* c, const void* keys, uint64_t n, void* vals, uint8_t* found,
                    void* stream) {
  if (!c || (!keys && n)) return fail(GAL_ERR_INVALID, "null argument");
  if (n == 0) return GAL_OK;
  return with_chain_kind(c->kind, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    k_chain_query<T><<<blocks_for(n), 256, 0, (hipStream_t)stream>>>(
        static_cast<T*>(c->dev), static_cast<const typename T::key_type*>(keys), n,
        static_cast<typename T::val_type*>(vals), found);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GAL_OK : hip_fail(e, "chain query");
  });
}

int gal_chain_stats(gal_chain* c, uint64_t* max_len, uint64_t* total_blocks, uint64_t* nblocks,
                    uint64_t* failed) {
  if (!c) return fail(GAL_ERR_INVALID, "null argument");
  if (hipDeviceSynchronize() != hipSuccess) return fail(GAL_ERR_HIP, "sync");
  return with_chain_kind(c->kind, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    T* dev = static_cast<T*>(c->dev);
    T host;
    if (hipMemcpy(&host, dev, sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
      return fail(GAL_ERR_HIP, "stats copy");
    dev->chain_stats(max_len, total_blocks);
    if (nblocks) *nblocks = host.nblocks;
    if (failed) *failed = host.failed;
    return hipDeviceSynchronize() == hipSuccess ? GAL_OK : fail(GAL_ERR_HIP, "stats");
  });
}

int gal_chain_lengths(gal_chain* c, uint32_t* out, void* stream) {
  if (!c || !out) return fail(GAL_ERR_INVALID, "null argument");
  return with_chain_kind(c->kind, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    gallatin::data_structs::chain_lengths_kernel<T>
        <<<blocks_for(c->nblocks), 256, 0, (hipStream_t)stream>>>(static_cast<T*>(c->dev), out,
                                                                  c->nblocks);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GAL_OK : hip_fail(e, "chain lengths");
  });
}

int gal_chain_print(gal_chain* c) {
  if (!c) return fail(GAL_ERR_INVALID, "null argument");
  return with_chain_kind(c->kind, [&](auto* tag) -> int {
    using T = std::remove_pointer_t<decltype(tag)>;
    static_cast<T*>(c->dev)->print_chain_stats();
    return GAL_OK;
  });
}

}  // extern "C"

// ---- block queues ----
//
// block_queue<uint64_t, 256> (the reference test's node size) and
// block_queue<uint64_t, 16> (small nodes: many node changes), by kind.

using gallatin::data_structs::block_queue;
using bqueue_256 = block_queue<uint64_t, 256>;
using bqueue_16 = block_queue<uint64_t, 16>;

struct gal_bqueue {
  int kind = 0;
  void* dev = nullptr;
};

namespace {

template <typename F>
int with_bqueue_kind(int kind, F&& f) {
  switch (kind) {
    case 0: return f(static_cast<bqueue_256*>(nullptr));
    case 1: return f(static_cast<bqueue_16*>(nullptr));
    default: return fail(GAL_ERR_INVALID, "unknown block queue kind");
  }
}

template <typename Q>
__global__ void k_bq_enqueue(Q* q, const uint64_t* items, uint64_t n, uint8_t* ok, int serial) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (serial) {  // one thread, in order
    if (i == 0)
      for (uint64_t k = 0; k < n; ++k) {
        const bool r = q->enqueue(items[k]);
        if (ok) ok[k] = r;
      }
    return;
  }
  if (i >= n) return;
  const bool r = q->enqueue(items[i]);
  if (ok) ok[i] = r;
}

template <typename Q>
__global__ void k_bq_dequeue(Q* q, uint64_t* out, uint8_t* ok, uint64_t n, int serial) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (serial) {
    if (i == 0)
      for (uint64_t k = 0; k < n; ++k) {
        uint64_t v = 0;
        const bool r = q->dequeue(v);
        out[k] = r ? v : 0;
        if (ok) ok[k] = r;
      }
    return;
  }
  if (i >= n) return;
  uint64_t v = 0;
  const bool r = q->dequeue(v);
  out[i] = r ? v : 0;
  if (ok) ok[i] = r;
}

// Nodes in the list, nodes on the retired list, refused enqueues.
template <typename Q>
__global__ void k_bq_stats(Q* q, uint64_t* out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  uint64_t live = 0, retired = 0;
  for (uint64_t a = q->head; a > Q::kAppending && live < Q::kHops; a = Q::as_node(a)->next)
    ++live;
  for (uint64_t a = q->retired; a && retired < Q::kHops; a = Q::as_node(a)->retired) ++retired;
  out[0] = live;
  out[1] = retired;
  out[2] = q->failed;
}

}  // namespace

extern "C" {

int gal_bqueue_create(int kind, gal_bqueue** out) {
  if (!out) return fail(GAL_ERR_INVALID, "null argument");
  if (!current_global()) return fail(GAL_ERR_INVALID, "no global allocator");
  return with_bqueue_kind(kind, [&](auto* tag) -> int {
    using Q = std::remove_pointer_t<decltype(tag)>;
    Q* dev = Q::generate_on_device();
    if (!dev) return fail(GAL_ERR_HIP, "block queue creation failed");
    *out = new gal_bqueue{kind, dev};
    return GAL_OK;
  });
}

int gal_bqueue_destroy(gal_bqueue* q) {
  if (!q) return GAL_OK;
  const int r = with_bqueue_kind(q->kind, [&](auto* tag) -> int {
    using Q = std::remove_pointer_t<decltype(tag)>;
    Q::free_on_device(static_cast<Q*>(q->dev));
    return GAL_OK;
  });
  delete q;
  return r;
}

int gal_bqueue_enqueue(gal_bqueue* q, const uint64_t* items, uint64_t n, uint8_t* ok,
                       int serial, void* stream) {
  if (!q || (!items && n)) return fail(GAL_ERR_INVALID, "null argument");
  if (n == 0) return GAL_OK;
  return with_bqueue_kind(q->kind, [&](auto* tag) -> int {
    using Q = std::remove_pointer_t<decltype(tag)>;
    k_bq_enqueue<Q><<<serial ? 1u : blocks_for(n), 256, 0, (hipStream_t)stream>>>(
        static_cast<Q*>(q->dev), items, n, ok, serial);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GAL_OK : hip_fail(e, "block queue enqueue");
  });
}

int gal_bqueue_dequeue(gal_bqueue* q, uint64_t* out, uint8_t* ok, uint64_t n, int serial,
                       void* stream) {
  if (!q || (!out && n)) return fail(GAL_ERR_INVALID, "null argument");
  if (n == 0) return GAL_OK;
  return with_bqueue_kind(q->kind, [&](auto* tag) -> int {
    using Q = std::remove_pointer_t<decltype(tag)>;
    k_bq_dequeue<Q><<<serial ? 1u : blocks_for(n), 256, 0, (hipStream_t)stream>>>(
        static_cast<Q*>(q->dev), out, ok, n, serial);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GAL_OK : hip_fail(e, "block queue dequeue");
  });
}

int gal_bqueue_reclaim(gal_bqueue* q, uint64_t* freed) {
  if (!q) return fail(GAL_ERR_INVALID, "null argument");
  if (hipDeviceSynchronize() != hipSuccess) return fail(GAL_ERR_HIP, "sync");
  return with_bqueue_kind(q->kind, [&](auto* tag) -> int {
    using Q = std::remove_pointer_t<decltype(tag)>;
    const uint64_t n = Q::reclaim_on_device(static_cast<Q*>(q->dev), false);
    if (freed) *freed = n;
    return hipDeviceSynchronize() == hipSuccess ? GAL_OK : fail(GAL_ERR_HIP, "reclaim");
  });
}

int gal_bqueue_stats(gal_bqueue* q, uint64_t* nodes, uint64_t* retired, uint64_t* failed) {
  if (!q) return fail(GAL_ERR_INVALID, "null argument");
  if (hipDeviceSynchronize() != hipSuccess) return fail(GAL_ERR_HIP, "sync");
  return with_bqueue_kind(q->kind, [&](auto* tag) -> int {
    using Q = std::remove_pointer_t<decltype(tag)>;
    uint64_t* d = nullptr;
    uint64_t h[3] = {0, 0, 0};
    hipError_t e = hipMalloc(&d, sizeof(h));
    if (e != hipSuccess) return hip_fail(e, "stats buffer");
    k_bq_stats<Q><<<1, 64>>>(static_cast<Q*>(q->dev), d);
    e = hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "stats copy");
    if (nodes) *nodes = h[0];
    if (retired) *retired = h[1];
    if (failed) *failed = h[2];
    return GAL_OK;
  });
}

}  // extern "C"

// ---- serving loop ----
//
// Device-resident sessions, each owning a chunked_vector of 64-bit words
// on the global allocator: chunks of 128 B doubling to 64 KiB, then further
// 64 KiB chunks (at most 512 of them, 32 MiB past the first 128 KiB).  A
// step is three launches on one stream:
//   1. session pass: every live session ends with probability
//      end_q32 / 2^32 (its vector is freed, its slot returned), or appends
//      tokens x 32 words (tokens drawn from a 1024-entry quantile table).
//      A wave serves 64 sessions: their missing chunks are requested by the
//      whole wave in one allocator call per round, then the wave writes the
//      sessions' words one session after another with 16-byte stores;
//   2. admission: the step's arrivals take free slots if the allocator's
//      free segments exceed `reserve_segments`, else all are refused;
//   3. table traffic: ht_ops operations on the extendible hash table, 70 %
//      insert / 30 % remove of keys uniform in [1, 2^28].
// Word p of session id holds (id << 32) | p.  An append is cut short only
// when a chunk cannot be allocated or the vector is at its 4.2 M-word
// limit (words refused).

using session_vec = gallatin::data_structs::chunked_vector<uint64_t, 16, 8192, 512>;

struct serving_session {
  session_vec vec;
  uint64_t id;
  uint32_t alive;
  uint32_t born;
};

enum {
  kSrvAdmitted = 0, kSrvRefused, kSrvEnded, kSrvWords, kSrvWordsRefused, kSrvChunks,
  kSrvChunkFailed, kSrvChunksFreed, kSrvHtInserted, kSrvHtRemoved, kSrvHtAbsent,
  kSrvHtFailed, kSrvCount = 16
};

struct gal_serving {
  uint32_t max_sessions = 0;
  uint64_t seed = 0;
  serving_session* sessions = nullptr;  // device
  uint32_t* free_slots = nullptr;       // device stack
  int32_t* free_top = nullptr;          // device
  uint32_t* queue = nullptr;            // device work queue of the session pass
  uint32_t* tokens = nullptr;           // device, 1024 entries
  uint64_t* free_segments = nullptr;    // device
  table_large* table = nullptr;
  // the table pass runs beside the session pass on a stream of its own
  // (default priority), forked from and joined back to the step's stream
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  bool sessions_ready = false;  // k_serve_init has run: sessions may own memory
};

namespace {

// Adds the number of active lanes with `pred` set to *c with one atomic per
// wave.  Every active lane must call it.  Per-lane atomics on one shared
// counter serialise at one memory channel (about 88 M/s on MI355X), which
// would dominate a kernel of a million short operations.
__device__ __forceinline__ void count_wave(uint64_t* c, bool pred) {
  const uint64_t m = __ballot(pred);
  if (m && gallatin::wave::is_first())
    gallatin::utils::atomic_add(c, (uint64_t)gallatin::wave::count(m));
}

__device__ __forceinline__ uint64_t mix3(uint64_t a, uint64_t b, uint64_t c) {
  uint64_t x = a * 0x9E3779B97F4A7C15ULL ^ b;
  x ^= x >> 31;
  x *= 0xBF58476D1CE4E5B9ULL;
  x ^= c + 0x94D049BB133111EBULL;
  x ^= x >> 29;
  x *= 0x94D049BB133111EBULL;
  return x ^ (x >> 32);
}

// The free-slot stack starts as slots 0, m, 2m, ... (mod n) from the top,
// m coprime to n and near 0.618 n.  In slot order, sessions admitted while
// arrivals outrun ends would take consecutive fresh slots, and the few
// waves owning those slots would do all the appends of that young cohort
// (their vectors grow for about 13 steps) while the others idle: those
// steps took 1.6-2.7 ms against 1.0 ms.  Spread out, every wave gets its
// share of young sessions.
__global__ void k_serve_init(serving_session* ss, uint32_t* free_slots, int32_t* top, uint32_t n,
                             uint64_t m) {
  const uint64_t i = gallatin::utils::get_tid();
  if (i < n) {
    ss[i].alive = 0;
    ss[i].id = ~0ULL;
    free_slots[i] = (uint32_t)(((uint64_t)(n - 1 - i) * m) % n);
  }
  if (i == 0) *top = (int32_t)n;
}

// Words [at, stop) of session `sid` into chunk memory `w` that starts at
// position cs: the `width` lanes of a group (this lane is number `sl` in
// it) write 16-byte pairs (chunks are at least 128-byte aligned and start at
// even positions).
__device__ __forceinline__ void write_words(uint64_t* w, uint64_t cs, uint64_t at, uint64_t stop,
                                            uint64_t sid, uint32_t sl, uint32_t width) {
  const uint64_t tag = sid << 32;
  if (at & 1) {  // odd head word
    if (sl == 0) w[at - cs] = tag | at;
    ++at;
  }
  const uint64_t pairs = (stop - at) >> 1;
  ulonglong2* v = reinterpret_cast<ulonglong2*>(w + (at - cs));
  for (uint64_t j = sl; j < pairs; j += width) {
    const uint64_t p = at + 2 * j;
    v[j] = make_ulonglong2(tag | p, tag | (p + 1));
  }
  if ((stop - at) & 1 && sl == 0) w[stop - 1 - cs] = tag | (stop - 1);
}

// Session pass: each wave takes 64 session slots at a time.
__global__ __launch_bounds__(256) void k_serve_sessions(serving_session* ss, uint32_t n,
                                                        uint32_t* free_slots, int32_t* top,
                                                        uint32_t* queue, const uint32_t* tokens,
                                                        uint64_t seed, uint32_t step,
                                                        uint32_t end_q32, uint64_t* ctr) {
  gal_state* g = global_gallatin;
  const uint32_t lane = gallatin::wave::lane();
  uint64_t ended = 0, words = 0, refused = 0, chunks = 0, chunk_fail = 0, freed = 0;
  // waves take 64 slots at a time, grid-stride (the launch has about one
  // wave per 64 slots, so a shared queue counter would only add atomics on
  // one address)
  const uint32_t wave_id = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  (void)queue;
  for (uint64_t base = (uint64_t)wave_id * 64; base < n; base += (uint64_t)nwaves * 64) {
    const uint32_t slot = (uint32_t)base + lane;
    bool alive = slot < n && ss[slot].alive;
    const uint64_t id = alive ? ss[slot].id : 0;
    const uint64_t r = mix3(seed, id, step);
    const bool ends = alive && (uint32_t)r < end_q32;
    // slots of ending sessions go back on the stack, one push per wave
    const uint64_t em = __ballot(ends);
    int32_t at0 = 0;
    if (em && gallatin::wave::is_first()) at0 = atomicAdd(top, (int32_t)gallatin::wave::count(em));
    at0 = (int32_t)gallatin::wave::readfirst((uint32_t)at0);
    // ending sessions' vectors are freed by the whole wave, one at a time
    for (uint64_t m = em; m; m &= m - 1) {
      const int l = gallatin::wave::first(m);
      session_vec* ev = reinterpret_cast<session_vec*>(
          gallatin::wave::readlane((uint64_t)&ss[slot].vec, l));
      const uint32_t nf = ev->free_all_wave();
      if ((int)lane == l) freed += nf;
    }
    if (ends) {  // return the slot
      ss[slot].alive = 0;
      free_slots[at0 + (int32_t)gallatin::wave::count(em & gallatin::wave::lanes_below())] = slot;
      ++ended;
    }
    // each lane: its session's words [at, end) of this step
    const bool live = alive && !ends;
    session_vec* v = live ? &ss[slot].vec : nullptr;
    const uint64_t at = live ? v->size : 0;
    const uint64_t want = live ? (uint64_t)tokens[(r >> 32) & 1023] * 32 : 0;
    uint64_t end = at + want;
    if (end > session_vec::max_items) end = session_vec::max_items;
    // missing chunks of [at, end): the wave requests one chunk (or the tail
    // index) per lane per round, all lanes in one allocator call
    uint32_t d = 0, dlast = 0;
    bool need = live && end > at;
    if (need) {
      d = session_vec::chunk_of(at);
      dlast = session_vec::chunk_of(end - 1);
      while (d <= dlast && v->chunk(d) != nullptr) ++d;
      need = d <= dlast;
    }
    while (__ballot(need)) {
      const bool idx = need && v->needs_index(d);
      const uint64_t bytes = idx ? session_vec::index_bytes
                                 : sizeof(uint64_t) * session_vec::chunk_items(d);
      void* p = gallatin::allocators::detail::malloc_impl(g, bytes, need);
      if (need) {
        if (!p) {  // this chunk cannot be had: the append stops before it
          ++chunk_fail;
          const uint64_t cs = session_vec::chunk_start(d);
          end = cs > at ? cs : at;
          need = false;
        } else if (idx) {
          v->set_index(static_cast<uint64_t**>(p));
        } else {
          v->set_chunk(d, static_cast<uint64_t*>(p));
          ++chunks;
          need = ++d <= dlast;
        }
      }
    }
    if (live) refused += at + want - end;
    // appends, four sessions at a time, each written by a group of 16 lanes
    const uint32_t grp = lane >> 4, sl = lane & 15;
    for (uint64_t todo = __ballot(live && end > at); todo;) {
      int pick = -1;
      for (uint32_t k = 0; k < 4 && todo; ++k, todo &= todo - 1)
        if (k == grp) pick = gallatin::wave::first(todo);
      // every lane takes part in the fetches (a lane reads a masked lane's
      // value as undefined); a group without a session copies its own lane
      const int src = pick < 0 ? (int)lane : pick;
      session_vec* sv = reinterpret_cast<session_vec*>(gallatin::wave::fetch((uint64_t)v, src));
      const uint64_t sid = gallatin::wave::fetch(id, src);
      uint64_t a = gallatin::wave::fetch(at, src);
      const uint64_t e_src = gallatin::wave::fetch(end, src);
      const uint64_t e = pick < 0 ? a : e_src;
      while (a < e) {
        const uint32_t c = session_vec::chunk_of(a);
        const uint64_t cs = session_vec::chunk_start(c);
        const uint64_t ce = cs + session_vec::chunk_items(c);
        const uint64_t stop = ce < e ? ce : e;
        write_words(sv->chunk(c), cs, a, stop, sid, sl, 16);
        a = stop;
      }
      if (pick >= 0 && sl == 0) sv->size = e;
    }
    if (live) words += end - at;
  }
  const uint64_t vals[6] = {ended, words, refused, chunks, chunk_fail, freed};
  const int idx[6] = {kSrvEnded, kSrvWords, kSrvWordsRefused, kSrvChunks, kSrvChunkFailed,
                      kSrvChunksFreed};
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    uint64_t t = vals[i];
    for (int off = 32; off >= 1; off >>= 1) t += gallatin::wave::fetch(t, (int)(lane ^ (uint32_t)off));
    if (lane == 0 && t) gallatin::utils::atomic_add(&ctr[idx[i]], t);
  }
}

// Free memory of the global allocator for the admission gate (one
// workgroup).  out[0]: free segments (bits set in the segment tree); out[1]:
// the free blocks of formatted segments, each weighted by its share of a
// segment, times blocks_per_segment (a block of class t holds 2^t / B of a
// segment), so out[0] + out[1] / B is the free memory in segments.
__global__ void k_serve_free_segments(uint64_t* out) {
  const gal_state* g = global_gallatin;
  const gallatin::allocators::gal_tree& t = g->seg_tree;
  const uint32_t L = t.num_layers - 1;
  uint64_t c = 0, w = 0;
  for (uint64_t i = threadIdx.x; i < t.layer_words[L]; i += blockDim.x)
    c += __popcll(gallatin::utils::atomic_load(t.words + t.layer_off[L] + i));
  for (uint64_t seg = threadIdx.x; seg < g->num_segments; seg += blockDim.x) {
    const uint64_t st = gallatin::utils::atomic_load(&g->seg_state[seg]);
    const uint32_t tree = (uint32_t)(st & 0xFFFF);
    const int32_t active = (int32_t)(st >> 32);
    if (tree < g->num_trees && active >= 0) w += (uint64_t)(active + 1) << tree;
  }
  for (int off = 32; off >= 1; off >>= 1) {
    c += gallatin::wave::fetch(c, (int)(gallatin::wave::lane() ^ (uint32_t)off));
    w += gallatin::wave::fetch(w, (int)(gallatin::wave::lane() ^ (uint32_t)off));
  }
  if (gallatin::wave::lane() == 0) {
    atomicAdd((unsigned long long*)&out[0], (unsigned long long)c);
    atomicAdd((unsigned long long*)&out[1], (unsigned long long)w);
  }
}

__global__ void k_serve_reset(uint32_t* queue, uint64_t* free_segments) {
  if (threadIdx.x == 0) {
    *queue = 0;
    free_segments[0] = 0;
    free_segments[1] = 0;
  }
}

// gate 0: admit while the segment tree holds more than `reserve` free
// segments; gate 1: while the free memory, whole free segments plus the free
// blocks of formatted segments counted in segments, exceeds `reserve`.
__global__ void k_serve_admit(serving_session* ss, uint32_t* free_slots, int32_t* top,
                              uint32_t arrivals, uint64_t reserve, uint32_t gate,
                              const uint64_t* free_segments, uint32_t step, uint64_t* ctr) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= arrivals) return;
  // One pop of the free-slot stack per wave: the wave takes as many slots
  // as it has arrivals, and the lanes past what the stack held give theirs
  // back.
  const uint64_t free_now =
      free_segments[0] + (gate ? free_segments[1] / global_gallatin->blocks_per_segment : 0);
  const bool room = free_now > reserve;
  const uint64_t want = __ballot(room);
  const uint32_t rank = gallatin::wave::count(want & gallatin::wave::lanes_below());
  int32_t top0 = 0;
  if (want && gallatin::wave::is_first()) {
    const int32_t n = (int32_t)gallatin::wave::count(want);
    top0 = atomicSub(top, n);
    const int32_t got = top0 < 0 ? 0 : (top0 < n ? top0 : n);
    if (got < n) atomicAdd(top, n - got);
  }
  top0 = (int32_t)gallatin::wave::readfirst((uint32_t)top0);
  const bool admitted = room && (int32_t)rank < top0;
  if (admitted) {
    const uint32_t slot = free_slots[top0 - 1 - (int32_t)rank];
    serving_session& se = ss[slot];
    se.id = ((uint64_t)step << 20) | i;
    se.born = step;
    se.vec.init();
    se.alive = 1;
  }
  count_wave(&ctr[kSrvAdmitted], admitted);
  count_wave(&ctr[kSrvRefused], !admitted);
}

__global__ void k_serve_table(table_large* t, uint64_t seed, uint32_t step, uint32_t n,
                              uint64_t* ctr) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t r = mix3(seed ^ 0x5eedULL, step, i);
  const uint64_t key = (r & ((1ULL << 28) - 1)) + 1;
  const bool ins = (r >> 40) % 10 < 7;
  const bool ok = ins ? t->insert(key, r) : t->remove(key);
  count_wave(&ctr[kSrvHtInserted], ins && ok);
  count_wave(&ctr[kSrvHtFailed], ins && !ok);
  count_wave(&ctr[kSrvHtRemoved], !ins && ok);
  count_wave(&ctr[kSrvHtAbsent], !ins && !ok);
}

// One wave per live session (grid-stride): every word read back.  out[0]
// bad words, [1] live sessions, [2] words, [3] chunk bytes held.
__global__ __launch_bounds__(256) void k_serve_check(serving_session* ss, uint32_t n,
                                                     unsigned long long* out) {
  const uint32_t lane = gallatin::wave::lane();
  const uint32_t wave_id = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  uint64_t bad = 0, live = 0, words = 0, held = 0;
  for (uint32_t i = wave_id; i < n; i += nwaves) {
    if (!ss[i].alive) continue;
    session_vec& v = ss[i].vec;
    const uint64_t tag = ss[i].id << 32;
    const uint64_t size = v.size;
    ++live;
    words += size;
    for (uint32_t c = 0; c < session_vec::n_chunks; ++c) {
      const uint64_t cs = session_vec::chunk_start(c);
      const uint64_t* w = v.chunk(c);
      if (w) held += session_vec::chunk_items(c) * sizeof(uint64_t);
      if (cs >= size) {
        if (!w) break;  // chunks are installed in order
        continue;
      }
      const uint64_t ce = cs + session_vec::chunk_items(c);
      const uint64_t stop = ce < size ? ce : size;
      if (!w) {
        bad += stop - cs;
        continue;
      }
      uint64_t b = 0;
      for (uint64_t p = cs + lane; p < stop; p += 64) b += w[p - cs] != (tag | p);
      for (int off = 32; off >= 1; off >>= 1) b += gallatin::wave::fetch(b, (int)(lane ^ (uint32_t)off));
      bad += b;
    }
    if (v.tail) held += session_vec::index_bytes;
  }
  if (lane == 0 && (bad | live | words | held)) {
    atomicAdd(&out[0], (unsigned long long)bad);
    atomicAdd(&out[1], (unsigned long long)live);
    atomicAdd(&out[2], (unsigned long long)words);
    atomicAdd(&out[3], (unsigned long long)held);
  }
}

__global__ void k_serve_free_all(serving_session* ss, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !ss[i].alive) return;
  ss[i].vec.free_all();
  ss[i].alive = 0;
}

}  // namespace

extern "C" {

int gal_serving_create(uint32_t max_sessions, uint64_t seed, const uint32_t* tokens_table,
                       gal_serving** out) {
  if (!out || !tokens_table || max_sessions == 0) return fail(GAL_ERR_INVALID, "bad arguments");
  if (!current_global()) return fail(GAL_ERR_INVALID, "no global allocator");
  gal_serving* sv = new gal_serving();
  sv->max_sessions = max_sessions;
  sv->seed = seed;
  hipError_t e = hipSuccess;
  if ((e = hipMalloc(&sv->sessions, sizeof(serving_session) * max_sessions)) != hipSuccess ||
      (e = hipMalloc(&sv->free_slots, sizeof(uint32_t) * max_sessions)) != hipSuccess ||
      (e = hipMalloc(&sv->free_top, sizeof(int32_t))) != hipSuccess ||
      (e = hipMalloc(&sv->queue, sizeof(uint32_t))) != hipSuccess ||
      (e = hipMalloc(&sv->free_segments, 2 * sizeof(uint64_t))) != hipSuccess ||
      (e = hipMalloc(&sv->tokens, sizeof(uint32_t) * 1024)) != hipSuccess ||
      (e = hipMemcpy(sv->tokens, tokens_table, sizeof(uint32_t) * 1024,
                     hipMemcpyHostToDevice)) != hipSuccess) {
    gal_serving_destroy(sv);
    return hip_fail(e, "serving buffers");
  }
  uint64_t m = ((uint64_t)max_sessions * 2654435769ULL >> 32) | 1;  // ~0.618 n
  while (std::gcd(m, (uint64_t)max_sessions) != 1) m += 2;
  if ((e = hipStreamCreateWithFlags(&sv->side, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&sv->fork, hipEventDisableTiming)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&sv->join, hipEventDisableTiming)) != hipSuccess) {
    gal_serving_destroy(sv);
    return hip_fail(e, "serving streams");
  }
  k_serve_init<<<(max_sessions + 255) / 256, 256>>>(sv->sessions, sv->free_slots, sv->free_top,
                                                    max_sessions, m % max_sessions ? m : 1);
  if ((e = hipDeviceSynchronize()) != hipSuccess) {
    gal_serving_destroy(sv);
    return hip_fail(e, "serving init");
  }
  sv->sessions_ready = true;
  sv->table = table_large::generate_on_device(seed);
  if (!sv->table) {
    gal_serving_destroy(sv);
    return fail(GAL_ERR_HIP, "serving table creation failed");
  }
  *out = sv;
  return GAL_OK;
}

int gal_serving_step(gal_serving* sv, uint32_t step, uint32_t arrivals, uint64_t reserve_segments,
                     uint32_t gate, uint32_t end_q32, uint32_t ht_ops, uint64_t* counters_dev,
                     void* stream) {
  if (!sv || !counters_dev) return fail(GAL_ERR_INVALID, "null argument");
  if (gate > 1) return fail(GAL_ERR_INVALID, "gate is 0 (free segments) or 1 (free memory)");
  if (arrivals >= (1u << 20)) return fail(GAL_ERR_INVALID, "at most 2^20 arrivals per step");
  hipStream_t st = (hipStream_t)stream;
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  // The table operations depend on nothing of the session pass: they run on
  // the side stream, after the earlier work of `st`, and `st` waits for
  // them at the end of the step.  The session pass is submitted first, so
  // that its workgroups are dispatched ahead of the table's.
  hipError_t e = hipSuccess;
  if (ht_ops && (e = hipEventRecord(sv->fork, st)) != hipSuccess)
    return hip_fail(e, "serving fork");
  k_serve_reset<<<1, 64, 0, st>>>(sv->queue, sv->free_segments);
  const uint32_t waves = (sv->max_sessions + 63) / 64;
  uint32_t groups = (uint32_t)cus * 8;
  if (groups * 4 > waves) groups = (waves + 3) / 4;
  k_serve_sessions<<<groups, 256, 0, st>>>(sv->sessions, sv->max_sessions, sv->free_slots,
                                           sv->free_top, sv->queue, sv->tokens, sv->seed, step,
                                           end_q32, counters_dev);
  if (ht_ops) {
    if ((e = hipStreamWaitEvent(sv->side, sv->fork, 0)) != hipSuccess)
      return hip_fail(e, "serving fork");
    k_serve_table<<<(ht_ops + 255) / 256, 256, 0, sv->side>>>(sv->table, sv->seed, step, ht_ops,
                                                              counters_dev);
    if ((e = hipEventRecord(sv->join, sv->side)) != hipSuccess)
      return hip_fail(e, "serving join");
  }
  k_serve_free_segments<<<1, 1024, 0, st>>>(sv->free_segments);
  if (arrivals)
    k_serve_admit<<<(arrivals + 255) / 256, 256, 0, st>>>(sv->sessions, sv->free_slots,
                                                          sv->free_top, arrivals, reserve_segments,
                                                          gate, sv->free_segments, step,
                                                          counters_dev);
  if (ht_ops && (e = hipStreamWaitEvent(st, sv->join, 0)) != hipSuccess)
    return hip_fail(e, "serving join");
  e = hipGetLastError();
  return e == hipSuccess ? GAL_OK : hip_fail(e, "serving step");
}

// Live sessions, words checked, mismatching words, chunk bytes held.
// Synchronous.
int gal_serving_check(gal_serving* sv, uint64_t* live, uint64_t* words, uint64_t* bad,
                      uint64_t* held_bytes) {
  if (!sv) return fail(GAL_ERR_INVALID, "null argument");
  unsigned long long* c = nullptr;
  hipError_t e = hipMalloc(&c, 4 * sizeof(unsigned long long));
  if (e != hipSuccess) return hip_fail(e, "check buffer");
  (void)hipMemset(c, 0, 4 * sizeof(unsigned long long));
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  k_serve_check<<<(unsigned)cus * 8, 256>>>(sv->sessions, sv->max_sessions, c);
  unsigned long long h[4] = {0, 0, 0, 0};
  e = hipMemcpy(h, c, sizeof(h), hipMemcpyDeviceToHost);
  (void)hipFree(c);
  if (e != hipSuccess) return hip_fail(e, "check copy");
  if (bad) *bad = h[0];
  if (live) *live = h[1];
  if (words) *words = h[2];
  if (held_bytes) *held_bytes = h[3];
  return GAL_OK;
}

int gal_serving_gate(gal_serving* sv, uint64_t* free_segments, uint64_t* free_memory_segments) {
  if (!sv) return fail(GAL_ERR_INVALID, "null argument");
  uint64_t h[2] = {0, 0};
  const hipError_t e = hipMemcpy(h, sv->free_segments, sizeof(h), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "gate copy");
  gal_allocator* g = gal_global_handle();
  gal_info info;
  if (!g || gal_get_info(g, &info) != GAL_OK) return fail(GAL_ERR_INVALID, "no global allocator");
  if (free_segments) *free_segments = h[0];
  if (free_memory_segments) *free_memory_segments = h[0] + h[1] / info.blocks_per_segment;
  return GAL_OK;
}

int gal_serving_table(gal_serving* sv, uint64_t* keys, uint64_t* buckets, uint64_t* failed) {
  if (!sv || !sv->table) return fail(GAL_ERR_INVALID, "null argument");
  gal_table t{1, sv->table};
  return gal_table_stats(&t, keys, buckets, nullptr, failed);
}

int gal_serving_destroy(gal_serving* sv) {
  if (!sv) return GAL_OK;
  if (sv->side) (void)hipStreamSynchronize(sv->side);
  if (sv->sessions && sv->sessions_ready) {  // before set-up, no session owns memory
    k_serve_free_all<<<(sv->max_sessions + 255) / 256, 256>>>(sv->sessions, sv->max_sessions);
    (void)hipDeviceSynchronize();
  }
  if (sv->table) table_large::free_on_device(sv->table);
  if (sv->sessions) (void)hipFree(sv->sessions);
  if (sv->free_slots) (void)hipFree(sv->free_slots);
  if (sv->free_top) (void)hipFree(sv->free_top);
  if (sv->queue) (void)hipFree(sv->queue);
  if (sv->free_segments) (void)hipFree(sv->free_segments);
  if (sv->tokens) (void)hipFree(sv->tokens);
  if (sv->fork) (void)hipEventDestroy(sv->fork);
  if (sv->join) (void)hipEventDestroy(sv->join);
  if (sv->side) (void)hipStreamDestroy(sv->side);
  delete sv;
  return GAL_OK;
}

}  // extern "C"

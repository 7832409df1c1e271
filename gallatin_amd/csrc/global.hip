// The library's global allocator instance and its bulk kernels
// (global_allocator.cuh:41-97), and the allocator-backed data structures
// that allocate through it (fixed_vector, extendible_hash_table).  This is
// the only translation unit of the library that includes
// global_allocator.hpp, which defines the device symbol global_gallatin.
#include <hip/hip_runtime.h>

#include "gallatin/allocators/global_allocator.hpp"
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
  uint64_t hdr[5] = {0, 0, 0, 0, 0};  // directory, level, seed, buckets, failed
  if (hipMemcpy(hdr, t->dev, sizeof(hdr), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(GAL_ERR_HIP, "stats copy");
  if (keys)
    *keys = t->kind == 0 ? table_small::count_keys(static_cast<table_small*>(t->dev))
                         : table_large::count_keys(static_cast<table_large*>(t->dev));
  if (buckets) *buckets = hdr[3];
  if (level) *level = hdr[1];
  if (failed) *failed = hdr[4];
  return GAL_OK;
}

}  // extern "C"

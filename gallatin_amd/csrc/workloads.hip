// Workload kernels behind the C ABI: bulk malloc/free, the single-wave
// deterministic stream, the plain fixed-size method, the mixed-size churn and
// the atomic / copy ceilings used for the roofline.
//
// Reference counterparts: tests/src/gallatin_test.cu:59-112
// (alloc_one_size_pointer / free_one_size_pointer) and
// tests/src/gallatin_churn.cu:37-95 (pointer_churn_kernel).
#include <hip/hip_runtime.h>

#include "gallatin/allocators/batched.hpp"
#include "gallatin/allocators/gallatin.hpp"
#include "gallatin/data_structs/queue.hpp"
#include "gallatin/streams.h"
#include "gallatin_capi.h"
#include "internal.hpp"

using namespace gallatin;
using namespace gallatin::allocators;
using namespace gallatin_capi;

namespace {

#include "lane_counts.hpp"

// Payload fill and verify, spread over the whole wave: the buffers of the
// participating lanes are cut into 16-byte chunks numbered buffer after
// buffer (lane order), and lane r handles chunks r, r + 64, r + 128, ...
// A chunk's owner is the last lane whose chunk prefix is <= the chunk index,
// found by a 6-step binary search over the prefixes (one per lane, read with
// ds_bpermute).  Needs all 64 lanes of the wave active; the chunk loops are
// wave-uniform (every lane takes part in every ds_bpermute, a disabled source
// lane would read as 0) and only the memory access itself is predicated.
//
// A buffer of w words (w = ceil(size / 8)) lives in a slice whose size is a
// multiple of 16 bytes and at least 8 * w, so for odd w the slice extends at
// least 8 bytes past the last word: every chunk is accessed as one aligned
// 16-byte vector, and the verify ignores the upper half of a buffer's last
// chunk when w is odd.  Accesses go through global-address-space pointers
// (global_load / global_store), so stores never hold up the LDS waits of the
// owner search, and the verify keeps kVerifyDepth loads per lane in flight.
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) u64x2 global_u64x2;

constexpr int kVerifyDepth = 8;

struct chunk_plan {
  uint32_t prefix;    // chunks owned by lower lanes
  uint32_t total;     // chunks of the whole wave
  uint32_t full_end;  // chunks below this index of this lane's buffer are full
  uint64_t adjusted;  // buffer address - 16 * prefix
};

__device__ inline chunk_plan plan_chunks(bool on, const void* p, uint64_t words) {
  const uint32_t n = on ? (uint32_t)((words + 1) / 2) : 0u;
  chunk_plan c;
  wave::group_scan(n, ~0ULL, c.prefix, c.total);
  c.full_end = c.prefix + (on ? (uint32_t)(words / 2) : 0u);
  c.adjusted = (uint64_t)p - 16ULL * c.prefix;
  return c;
}

__device__ __forceinline__ int chunk_owner(uint32_t prefix, uint32_t c) {
  int l = 0;
  for (int step = 32; step; step >>= 1) {
    const uint32_t p = wave::fetch(prefix, l + step);
    if (p <= c) l += step;
  }
  return l;
}

__device__ inline void wave_fill(bool on, uint64_t* p, uint64_t words, uint64_t pat) {
  const chunk_plan cp = plan_chunks(on, p, words);
  const uint32_t lane = wave::lane();
  for (uint32_t base = 0; base < cp.total; base += 64) {
    const uint32_t c = base + lane;
    const bool in = c < cp.total;
    const int l = chunk_owner(cp.prefix, in ? c : cp.total - 1);
    const uint64_t a = wave::fetch(cp.adjusted, l) + 16ULL * c;
    const uint64_t v = wave::fetch(pat, l);
    if (in) *reinterpret_cast<global_u64x2*>(a) = u64x2{v, v};
  }
}

// Returns this lane's share of the mismatching words.
__device__ inline uint32_t wave_verify(bool on, const uint64_t* p, uint64_t words,
                                       uint64_t pat) {
  const chunk_plan cp = plan_chunks(on, p, words);
  const uint32_t lane = wave::lane();
  uint32_t bad = 0;
  for (uint32_t base = 0; base < cp.total; base += 64 * kVerifyDepth) {
    uint64_t a[kVerifyDepth], v[kVerifyDepth];
    bool in[kVerifyDepth], full[kVerifyDepth];
#pragma unroll
    for (int k = 0; k < kVerifyDepth; ++k) {
      const uint32_t c = base + 64u * k + lane;
      in[k] = c < cp.total;
      const int l = chunk_owner(cp.prefix, in[k] ? c : cp.total - 1);
      a[k] = wave::fetch(cp.adjusted, l) + 16ULL * c;
      v[k] = wave::fetch(pat, l);
      full[k] = c < wave::fetch(cp.full_end, l);
    }
    u64x2 x[kVerifyDepth];
#pragma unroll
    for (int k = 0; k < kVerifyDepth; ++k)
      if (in[k]) x[k] = *reinterpret_cast<const global_u64x2*>(a[k]);
#pragma unroll
    for (int k = 0; k < kVerifyDepth; ++k)
      if (in[k]) bad += (x[k].x != v[k]) + (full[k] && x[k].y != v[k]);
  }
  return bad;
}

}  // namespace

__global__ void k_malloc_bulk(gal_state* s, const uint64_t* sizes, uint64_t n, void** out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = detail::malloc_impl(s, sizes[i]);
}

__global__ void k_free_bulk(gal_state* s, void* const* ptrs, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  void* p = ptrs[i];
  if (p) detail::free_impl(s, p);
}

// One wave replays a stream of wave steps in order.
__global__ void k_wave_stream(gal_state* s, const gal_wave_step* steps, uint64_t nsteps,
                              uint64_t* results) {
  const uint32_t lane = threadIdx.x;
  const uint64_t base = (uint64_t)s->memory;
  for (uint64_t i = 0; i < nsteps; ++i) {
    const uint32_t kind = steps[i].kind;
    const uint64_t mask = steps[i].mask;
    const bool on = (mask >> lane) & 1;
    uint64_t res = kFail;
    if (kind == 0) {
      if (on) {
        void* p = detail::malloc_impl(s, steps[i].arg[lane]);
        if (p) res = (uint64_t)p - base;
      }
    } else if (on) {
      const uint64_t ref = steps[i].arg[lane];
      const uint64_t off = ref < i * 64 ? results[ref] : kFail;  // earlier steps only
      detail::free_impl(s, off == kFail ? nullptr : (void*)(base + off));
    }
    results[i * 64 + lane] = res;
    __threadfence_block();
  }
}

// Plain method: malloc(size), write the thread id, read it back, free.  An
// instance created with GAL_FLAG_NO_PAYLOAD_FENCE leaves the ordering of
// payload to the caller, so there the id is written and read with
// device-coherent accesses (they do not stay in one XCD's L2).
__global__ void k_plain(gal_state* s, uint64_t n, uint64_t size, uint64_t* ctr) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool coherent = (s->flags & kFlagNoPayloadFence) != 0;
  lane_counts c;
  if (tid < n) {
    uint32_t* p = static_cast<uint32_t*>(detail::malloc_impl(s, size));
    c.mallocs = 1;
    if (!p) {
      c.misses = 1;
    } else {
      if (coherent) {
        utils::atomic_store(p, (uint32_t)tid);
        c.bad = utils::atomic_load(p) != (uint32_t)tid;
      } else {
        volatile uint32_t* v = p;
        v[0] = (uint32_t)tid;
        c.bad = v[0] != (uint32_t)tid;
      }
      c.bytes = 8;
      detail::free_impl(s, (void*)p);
      c.pairs = 1;
    }
  }
  add_counters(ctr, c);
}

// Mixed-size churn: one live allocation per thread, replaced in round 0, in
// every later round whose coin is set, and freed after the last round.
__global__ void k_churn(gal_state* s, uint64_t seed, uint64_t n, uint32_t rounds,
                        uint64_t* ctr) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = tid < n;
  uint64_t* live = nullptr;
  uint32_t live_words = 0;
  uint32_t live_round = 0;
  lane_counts c;
  for (uint32_t r = 0; r <= rounds; ++r) {
    const bool replace =
        active && (r == 0 || r == rounds || gal_churn_coin(seed, tid, r) != 0);
    const bool do_free = replace && live != nullptr;
    c.bad += wave_verify(do_free, live, live_words, gal_churn_pattern(tid, live_round));
    if (do_free) {
      detail::free_impl(s, live);
      c.pairs += 1;
      c.bytes += (uint64_t)live_words * 8;
      live = nullptr;
    }
    if (r == rounds) break;
    const uint32_t size = (uint32_t)gal_churn_size(seed, tid, r);
    uint64_t* np = nullptr;
    if (replace) {
      np = static_cast<uint64_t*>(detail::malloc_impl(s, size));
      c.mallocs += 1;
      if (!np) c.misses += 1;
    }
    const bool filled = replace && np != nullptr;
    if (filled) {
      live = np;
      live_words = (uint32_t)gal_fill_words(size);
      live_round = r;
      c.bytes += (uint64_t)live_words * 8;
    }
    wave_fill(filled, np, gal_fill_words(size), gal_churn_pattern(tid, r));
  }
  add_counters(ctr, c);
}

// Block and multi-segment batch, allocation half: request i of the step
// mallocs its size and tags its first word.
__global__ void k_big_malloc(gal_state* s, uint64_t seed, uint32_t step, uint64_t n,
                             uint64_t n_small, void** ptrs, uint64_t* ctr) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  lane_counts c;
  if (i < n) {
    uint64_t* p = static_cast<uint64_t*>(
        detail::malloc_impl(s, gal_big_size(seed, step, i, n_small)));
    c.mallocs = 1;
    if (p) {
      p[0] = gal_big_tag(step, i);
      c.bytes = 8;
    } else {
      c.misses = 1;
    }
    ptrs[i] = p;
  }
  add_counters(ctr, c);
}

// Free half: the j-th thread checks and frees request gal_big_free_order(j).
__global__ void k_big_free(gal_state* s, uint64_t seed, uint32_t step, uint64_t n,
                           void* const* ptrs, uint64_t* ctr) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  lane_counts c;
  if (j < n) {
    const uint64_t i = gal_big_free_order(seed, step, j, n);
    uint64_t* p = static_cast<uint64_t*>(ptrs[i]);
    if (p) {
      c.bad = p[0] != gal_big_tag(step, i);
      c.bytes = 8;
      detail::free_impl(s, p);
      c.pairs = 1;
    }
  }
  add_counters(ctr, c);
}

// The same churn through the workgroup-batched path: a persistent grid walks
// chunks of blockDim stream threads; for every round the workgroup frees its
// replaced allocations and allocates the new ones as one batch step.
template <int kThreads>
__global__ __launch_bounds__(kThreads) void k_churn_batched(gal_state* s, uint64_t seed,
                                                            uint64_t n, uint32_t rounds,
                                                            uint64_t* ctr) {
  __shared__ batch::lds_state lds;
  batch::batch_init(lds);
  lane_counts c;
  const uint64_t nchunks = (n + blockDim.x - 1) / blockDim.x;
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t tid = chunk * blockDim.x + threadIdx.x;
    const bool active = tid < n;
    uint64_t* live = nullptr;
    uint32_t live_words = 0;
    uint32_t live_round = 0;
    for (uint32_t r = 0; r <= rounds; ++r) {
      const bool replace =
          active && (r == 0 || r == rounds || gal_churn_coin(seed, tid, r) != 0);
      const bool do_free = replace && live != nullptr;
      c.bad += wave_verify(do_free, live, live_words, gal_churn_pattern(tid, live_round));
      void* old = do_free ? live : nullptr;
      if (do_free) {
        c.pairs += 1;
        c.bytes += (uint64_t)live_words * 8;
        live = nullptr;
      }
      if (r == rounds) {
        batch::batch_free(s, old, lds);
        break;
      }
      const uint32_t size = (uint32_t)gal_churn_size(seed, tid, r);
      uint64_t* np = static_cast<uint64_t*>(batch::batch_step(s, old, replace, size, lds));
      if (replace) {
        c.mallocs += 1;
        if (!np) c.misses += 1;
      }
      const bool filled = replace && np != nullptr;
      if (filled) {
        live = np;
        live_words = (uint32_t)gal_fill_words(size);
        live_round = r;
        c.bytes += (uint64_t)live_words * 8;
      }
      wave_fill(filled, np, gal_fill_words(size), gal_churn_pattern(tid, r));
    }
  }
  add_counters(ctr, c);
}

// Atomic-rate probe: op i adds 1 to word ((i / lanes_per_address) * stride)
// % span and keeps the returned value live, as the allocator's claims do.
__global__ void k_bench_atomic(uint32_t* words, uint64_t span, uint64_t n_ops,
                               uint32_t lanes_per_address, uint32_t stride) {
  const uint64_t nthreads = (uint64_t)gridDim.x * blockDim.x;
  uint32_t acc = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_ops; i += nthreads) {
    const uint64_t idx = ((i / lanes_per_address) * stride) % span;
    acc += utils::atomic_add(&words[idx], 1u);
  }
  if (acc == 0xFFFFFFFFu) words[0] = acc;
}

// Tree-search probe: every wave runs `searches` wave-cooperative find_ge
// calls from pseudo-random starts, alternating between the instance's
// segment tree and class tree `tree` of partition 0.  Read only: the trees
// do not change, so counters over this kernel see the allocator's tree
// reads alone.
__global__ void k_bench_tree(const gal_state* s, uint32_t tree, uint32_t searches,
                             uint64_t* sink) {
  const uint64_t wid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const gal_tree& seg = s->seg_tree;
  const gal_tree& cls = s->sub_trees[tree * s->num_parts];
  uint64_t acc = 0;
  for (uint32_t i = 0; i < searches; ++i) {
    const gal_tree& t = (i & 1) ? cls : seg;
    uint64_t h = (wid * 0x9E3779B97F4A7C15ULL) ^ ((uint64_t)i * 0xC2B2AE3D27D4EB4FULL);
    h ^= h >> 29;
    acc += veb::find_ge(t, h % t.universe);
  }
  if (acc == 0x5DEECE66DULL) sink[0] = acc;
}

// Each workgroup copies whole tiles of kCopyVecs * blockDim.x 16-byte
// vectors (a contiguous 128 KiB at 1024 threads), tiles dealt round robin;
// the vectors past the last whole tile go grid-stride.  Contiguous
// per-workgroup tiles measured about 12 % faster than a pure grid-stride
// copy on MI355X.  Loads and stores are non-temporal: a sweep of 256, 512
// and 1024 threads x 4, 8 and 16 vectors found this shape the fastest,
// 6.0 TB/s against 5.2 TB/s for 512 threads with ordinary accesses.
constexpr int kCopyVecs = 8;
constexpr int kCopyThreads = 1024;

__global__ void __launch_bounds__(kCopyThreads)
    k_bench_copy(const u64x2* src_, u64x2* dst_, uint64_t n) {
  const global_u64x2* src = (const global_u64x2*)src_;
  global_u64x2* dst = (global_u64x2*)dst_;
  const uint64_t tile = (uint64_t)blockDim.x * kCopyVecs;
  const uint64_t tiles = n / tile;
  for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const uint64_t b = t * tile + threadIdx.x;
    u64x2 v[kCopyVecs];
#pragma unroll
    for (int u = 0; u < kCopyVecs; ++u)
      v[u] = __builtin_nontemporal_load(&src[b + (uint64_t)u * blockDim.x]);
#pragma unroll
    for (int u = 0; u < kCopyVecs; ++u)
      __builtin_nontemporal_store(v[u], &dst[b + (uint64_t)u * blockDim.x]);
  }
  const uint64_t nthreads = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = tiles * tile + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += nthreads)
    dst[i] = src[i];
}

// Scattered-piece copy: the buffers are cut into 2^k pieces of
// piece_vecs 16-byte vectors, and logical piece i is read from source piece
// (i * ma) mod 2^k and written to destination piece (i * mb) mod 2^k (odd
// multipliers, so both maps are bijections and every access is in bounds).
// Each wave walks 64 * kCopyVecs consecutive logical vectors per step, the
// same per-wave order as the churn's fill and verify, so this is the
// bandwidth ceiling of that access pattern at a given allocation size
// (non-temporal loads and stores, as the contiguous copy).
__global__ void __launch_bounds__(256)
    k_bench_scatter(const u64x2* src_, u64x2* dst_, uint64_t pieces, uint32_t piece_vecs,
                    uint64_t ma, uint64_t mb) {
  const global_u64x2* src = (const global_u64x2*)src_;
  global_u64x2* dst = (global_u64x2*)dst_;
  const uint64_t mask = pieces - 1;
  const uint64_t total = pieces * piece_vecs;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint32_t lane = threadIdx.x & 63;
  for (uint64_t base = wave * 64 * kCopyVecs; base < total; base += nwaves * 64 * kCopyVecs) {
    u64x2 v[kCopyVecs];
    uint64_t to[kCopyVecs];
#pragma unroll
    for (int k = 0; k < kCopyVecs; ++k) {
      const uint64_t x = base + (uint64_t)k * 64 + lane;
      const uint64_t piece = x / piece_vecs, off = x % piece_vecs;
      to[k] = ((piece * mb) & mask) * piece_vecs + off;
      if (x < total) v[k] = __builtin_nontemporal_load(&src[((piece * ma) & mask) * piece_vecs + off]);
    }
#pragma unroll
    for (int k = 0; k < kCopyVecs; ++k)
      if (base + (uint64_t)k * 64 + lane < total) __builtin_nontemporal_store(v[k], &dst[to[k]]);
  }
}

template <int kThreads>
static int launch_churn_batched(gal_allocator* a, uint64_t seed, uint64_t n, uint32_t rounds,
                                uint64_t* ctr, uint32_t workgroups, hipStream_t stream) {
  if (workgroups == 0) {
    int dev = 0, cus = 256, per_cu = 1;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_churn_batched<kThreads>,
                                                     kThreads, 0) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    workgroups = (uint32_t)(cus * per_cu);
  }
  const uint64_t chunks = (n + kThreads - 1) / kThreads;
  if (workgroups > chunks) workgroups = (uint32_t)chunks;
  k_churn_batched<kThreads><<<workgroups, kThreads, 0, stream>>>(a->dev, seed, n, rounds, ctr);
  return launch_status("churn_batched");
}

extern "C" {

int gal_malloc_bulk(gal_allocator* a, const uint64_t* sizes, uint64_t n, void** out,
                    uint32_t threads, void* stream) {
  if (!a || (!sizes && n) || (!out && n)) return fail(GAL_ERR_INVALID, "null argument");
  if (n == 0) return GAL_OK;
  if (threads == 0 || threads > 1024 || threads % 64) threads = 256;
  k_malloc_bulk<<<grid_for(n, threads), threads, 0, (hipStream_t)stream>>>(a->dev, sizes, n, out);
  return launch_status("malloc_bulk");
}

int gal_free_bulk(gal_allocator* a, void* const* ptrs, uint64_t n, uint32_t threads,
                  void* stream) {
  if (!a || (!ptrs && n)) return fail(GAL_ERR_INVALID, "null argument");
  if (n == 0) return GAL_OK;
  if (threads == 0 || threads > 1024 || threads % 64) threads = 256;
  k_free_bulk<<<grid_for(n, threads), threads, 0, (hipStream_t)stream>>>(a->dev, ptrs, n);
  return launch_status("free_bulk");
}

int gal_run_wave_stream(gal_allocator* a, const gal_wave_step* steps, uint64_t nsteps,
                        uint64_t* results, void* stream) {
  if (!a || (!steps && nsteps) || (!results && nsteps)) return fail(GAL_ERR_INVALID, "null argument");
  if (nsteps == 0) return GAL_OK;
  k_wave_stream<<<1, 64, 0, (hipStream_t)stream>>>(a->dev, steps, nsteps, results);
  return launch_status("wave_stream");
}

int gal_run_plain(gal_allocator* a, uint64_t n, uint64_t size, uint64_t* ctr,
                  uint32_t threads, void* stream) {
  if (!a || !ctr) return fail(GAL_ERR_INVALID, "null argument");
  if (n == 0) return GAL_OK;
  if (threads == 0 || threads > 1024 || threads % 64) threads = 256;
  k_plain<<<grid_for(n, threads), threads, 0, (hipStream_t)stream>>>(a->dev, n, size, ctr);
  return launch_status("plain");
}

int gal_run_churn(gal_allocator* a, uint64_t seed, uint64_t n, uint32_t rounds,
                  uint64_t* ctr, uint32_t threads, void* stream) {
  if (!a || !ctr) return fail(GAL_ERR_INVALID, "null argument");
  if (n == 0) return GAL_OK;
  if (threads == 0 || threads > 1024 || threads % 64) threads = 256;
  k_churn<<<grid_for(n, threads), threads, 0, (hipStream_t)stream>>>(a->dev, seed, n, rounds, ctr);
  return launch_status("churn");
}

int gal_run_big(gal_allocator* a, uint64_t seed, uint32_t step, uint64_t n_small,
                uint64_t n_large, void** ptrs, uint64_t* ctr, void* stream) {
  if (!a || !ctr || !ptrs) return fail(GAL_ERR_INVALID, "null argument");
  const uint64_t n = n_small + n_large;
  if (n == 0) return GAL_OK;
  if (n > (1ULL << 32)) return fail(GAL_ERR_INVALID, "at most 2^32 requests per step");
  hipStream_t st = (hipStream_t)stream;
  k_big_malloc<<<grid_for(n, 256), 256, 0, st>>>(a->dev, seed, step, n, n_small, ptrs, ctr);
  int rc = launch_status("big_malloc");
  if (rc != GAL_OK) return rc;
  k_big_free<<<grid_for(n, 256), 256, 0, st>>>(a->dev, seed, step, n, ptrs, ctr);
  return launch_status("big_free");
}

int gal_run_churn_batched(gal_allocator* a, uint64_t seed, uint64_t n, uint32_t rounds,
                          uint64_t* ctr, uint32_t workgroups, uint32_t threads,
                          void* stream) {
  if (!a || !ctr) return fail(GAL_ERR_INVALID, "null argument");
  if (n == 0) return GAL_OK;
  hipStream_t st = (hipStream_t)stream;
  switch (threads) {
    case 256: return launch_churn_batched<256>(a, seed, n, rounds, ctr, workgroups, st);
    case 512: return launch_churn_batched<512>(a, seed, n, rounds, ctr, workgroups, st);
    case 0:
    case 1024: return launch_churn_batched<1024>(a, seed, n, rounds, ctr, workgroups, st);
    default: return fail(GAL_ERR_INVALID, "batched churn runs 256, 512 or 1024 threads per workgroup");
  }
}

int gal_bench_atomic(uint32_t* words, uint64_t span, uint64_t n_ops,
                     uint32_t lanes_per_address, uint32_t stride, void* stream) {
  if (!words || span == 0 || lanes_per_address == 0 || stride == 0)
    return fail(GAL_ERR_INVALID, "bad atomic probe arguments");
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  k_bench_atomic<<<(unsigned)cus * 8, 256, 0, (hipStream_t)stream>>>(words, span, n_ops,
                                                                      lanes_per_address, stride);
  return launch_status("bench_atomic");
}

int gal_bench_copy(const void* src, void* dst, uint64_t bytes, void* stream) {
  if (!src || !dst || bytes % 16) return fail(GAL_ERR_INVALID, "bad copy arguments");
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  k_bench_copy<<<(unsigned)cus * 8, kCopyThreads, 0, (hipStream_t)stream>>>(
      static_cast<const u64x2*>(src), static_cast<u64x2*>(dst), bytes / 16);
  return launch_status("bench_copy");
}

int gal_bench_tree(gal_allocator* a, uint32_t tree, uint32_t searches, uint32_t waves,
                   uint64_t* sink_dev, void* stream) {
  if (!a || !sink_dev || waves == 0) return fail(GAL_ERR_INVALID, "bad tree probe arguments");
  if (tree >= a->host.num_trees) return fail(GAL_ERR_INVALID, "no such class tree");
  const unsigned blocks = (waves + 3) / 4;  // four waves per workgroup
  k_bench_tree<<<blocks, 256, 0, (hipStream_t)stream>>>(a->dev, tree, searches, sink_dev);
  return launch_status("bench_tree");
}

int gal_bench_scatter(const void* src, void* dst, uint64_t bytes, uint64_t piece_bytes,
                      void* stream) {
  const auto pow2 = [](uint64_t x) { return x && !(x & (x - 1)); };
  if (!src || !dst || !pow2(bytes) || !pow2(piece_bytes) || piece_bytes < 16 ||
      piece_bytes > bytes || (bytes / piece_bytes) < 2)
    return fail(GAL_ERR_INVALID, "bad scatter arguments");
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  k_bench_scatter<<<(unsigned)cus * 8, 256, 0, (hipStream_t)stream>>>(
      static_cast<const u64x2*>(src), static_cast<u64x2*>(dst), bytes / piece_bytes,
      (uint32_t)(piece_bytes / 16), 0x9E3779B97F4A7C15ULL, 0xC2B2AE3D27D4EB4FULL);
  return launch_status("bench_scatter");
}

}  // extern "C"

// ---- allocator-backed queues ----
//
// The device methods of Gallatin<> read the instance's geometry at run time,
// so one instantiation serves every gal_allocator.
using queue_alloc = Gallatin<16ULL << 20, 16, 4096>;
using queue_one = gallatin::data_structs::queue<uint64_t, queue_alloc>;
using queue_ring = gallatin::data_structs::circular_CAS_queue<uint64_t, 8, queue_alloc>;

struct gal_queue {
  int ring = 0;
  void* dev = nullptr;
};

namespace {

template <typename Q>
__global__ void k_queue_enqueue(Q* q, const uint64_t* items, uint64_t n, uint8_t* ok,
                                int serial) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (serial) {  // one thread, in order
    if (i == 0)
      for (uint64_t k = 0; k < n; ++k) {
        const bool r = q->try_enqueue(items[k]);
        if (ok) ok[k] = r;
      }
    return;
  }
  if (i >= n) return;
  const bool r = q->try_enqueue(items[i]);
  if (ok) ok[i] = r;
}

template <typename Q>
__global__ void k_queue_dequeue(Q* q, uint64_t* out, uint8_t* ok, uint64_t n, int serial) {
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

}  // namespace

extern "C" {

int gal_queue_create(gal_allocator* a, int ring, gal_queue** out) {
  if (!a || !out) return fail(GAL_ERR_INVALID, "null argument");
  queue_alloc* al = static_cast<queue_alloc*>(static_cast<gal_state*>(a->dev));
  void* dev = ring ? (void*)queue_ring::generate_on_device(al)
                   : (void*)queue_one::generate_on_device(al);
  if (!dev) return fail(GAL_ERR_HIP, "queue creation failed");
  *out = new gal_queue{ring ? 1 : 0, dev};
  return GAL_OK;
}

int gal_queue_destroy(gal_queue* q) {
  if (!q) return GAL_OK;
  if (q->ring)
    queue_ring::free_on_device(static_cast<queue_ring*>(q->dev));
  else
    queue_one::free_on_device(static_cast<queue_one*>(q->dev));
  delete q;
  return GAL_OK;
}

int gal_queue_enqueue(gal_queue* q, const uint64_t* items, uint64_t n, uint8_t* ok, int serial,
                      void* stream) {
  if (!q || (!items && n)) return fail(GAL_ERR_INVALID, "null argument");
  if (n == 0) return GAL_OK;
  const unsigned g = serial ? 1u : grid_for(n, 256);
  if (q->ring)
    k_queue_enqueue<<<g, 256, 0, (hipStream_t)stream>>>(static_cast<queue_ring*>(q->dev), items,
                                                        n, ok, serial);
  else
    k_queue_enqueue<<<g, 256, 0, (hipStream_t)stream>>>(static_cast<queue_one*>(q->dev), items, n,
                                                        ok, serial);
  return launch_status("queue enqueue");
}

int gal_queue_dequeue(gal_queue* q, uint64_t* out, uint8_t* ok, uint64_t n, int serial,
                      void* stream) {
  if (!q || (!out && n)) return fail(GAL_ERR_INVALID, "null argument");
  if (n == 0) return GAL_OK;
  const unsigned g = serial ? 1u : grid_for(n, 256);
  if (q->ring)
    k_queue_dequeue<<<g, 256, 0, (hipStream_t)stream>>>(static_cast<queue_ring*>(q->dev), out, ok,
                                                        n, serial);
  else
    k_queue_dequeue<<<g, 256, 0, (hipStream_t)stream>>>(static_cast<queue_one*>(q->dev), out, ok,
                                                        n, serial);
  return launch_status("queue dequeue");
}

}  // extern "C"

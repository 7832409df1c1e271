// Batched request streams: batches of mixed malloc and free requests, one
// persistent launch per batch, with the launches of a whole stream captured
// in one hipGraph and replayed.
//
// An op is a malloc of `arg` bytes whose pointer goes to ptrs[slot], or a
// free of the pointer in ptrs[slot] (written by an earlier batch; nullptr,
// from a malloc that failed, is skipped).  Every allocation holds its slot
// number in its first 8 bytes, checked when it is freed.  After the last
// batch a cleanup launch frees every pointer still in the table, so a
// replay starts from an empty table.
//
// Two ways to serve a batch: the workgroup-batched path (batch_step of
// batched.hpp) and the plain per-thread malloc / free of gallatin.hpp.  No
// reference counterpart; the reference's tests drive the allocator with one
// kernel per operation type (tests/src/gallatin_test.cu:59-112).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "gallatin/allocators/batched.hpp"
#include "gallatin/allocators/slice_cache.hpp"
#include "gallatin/allocators/gallatin.hpp"
#include "gallatin_capi.h"
#include "internal.hpp"

using namespace gallatin;
using namespace gallatin::allocators;
using namespace gallatin_capi;

namespace {

__device__ inline uint64_t wave_total(uint64_t v) {
  for (int off = 32; off >= 1; off >>= 1)
    v += wave::fetch(v, (int)(wave::lane() ^ (uint32_t)off));
  return v;
}

struct op_counts {
  uint64_t pairs = 0, misses = 0, bad = 0, bytes = 0, mallocs = 0, skipped = 0;
  uint64_t polls = 0;  // polls of frees that found their entry unpublished
  // wall-clock ticks (wall_clock64, a scalar clock: the same on every lane of
  // a wave) the wave spent in malloc rounds, waiting for entries and
  // freeing, and its clock at kernel entry; only k_stream keeps them
  uint64_t malloc_ticks = 0, wait_ticks = 0, free_ticks = 0, start = wall_clock64();
};

// One atomic per counter per wave (every lane of the wave active).  The
// wave's time split is added once per wave (lane 0's clock readings), with
// its lifetime when `timed`.
__device__ inline void flush_counts(uint64_t* ctr, const op_counts& c, bool timed = false) {
  const uint64_t v[5] = {wave_total(c.pairs), wave_total(c.misses), wave_total(c.bad),
                         wave_total(c.bytes), wave_total(c.mallocs)};
  const uint64_t sk = wave_total(c.skipped);
  const uint64_t polls = wave_total(c.polls);
  const uint64_t life = wall_clock64() - c.start;
  if (wave::lane() == 0) {
#pragma unroll
    for (int i = 0; i < 5; ++i)
      if (v[i]) utils::atomic_add(&ctr[i], v[i]);
    if (sk) utils::atomic_add(&ctr[GAL_CTR_FREES_SKIPPED], sk);
    if (polls) utils::atomic_add(&ctr[GAL_CTR_WAIT_POLLS], polls);
    if (timed) {
      const uint64_t t[4] = {c.wait_ticks, life, c.malloc_ticks, c.free_ticks};
      const int at[4] = {GAL_CTR_WAIT_TICKS, GAL_CTR_WAVE_TICKS, GAL_CTR_MALLOC_TICKS,
                         GAL_CTR_FREE_TICKS};
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (t[i]) utils::atomic_add(&ctr[at[i]], t[i]);
    }
  }
}

// Pointer table protocol.  An entry is 0 until its malloc has run, then
// the pointer, or kMallocFailed when the malloc returned nullptr.  A free
// waits (bounded) for its entry to leave 0, so it never depends on launch
// boundaries: batches may run in one launch.  Entries and tags cross
// workgroups and XCDs through agent-scope atomic loads and stores (coherent
// at the device, not held in a non-coherent cache), the tag store is
// acknowledged before the entry is published, and the free reads the tag
// only after it has seen the entry.
constexpr uint64_t kMallocFailed = 1;

// Decode op i without waiting: a malloc (want, size) or a free (is_free)
// of table entry `slot`.
__device__ inline void decode_op(const gal_batch_op* ops, uint64_t i, bool active,
                                 uint64_t nslots, op_counts& c, bool& want, uint64_t& size,
                                 uint32_t& slot, bool& is_free) {
  want = false;
  is_free = false;
  size = 0;
  slot = 0;
  if (!active) return;
  const gal_batch_op op = ops[i];
  if (op.slot >= nslots || op.kind > 1) {  // not a valid op: counted, skipped
    c.bad += 1;
    return;
  }
  slot = op.slot;
  if (op.kind == 0) {
    want = true;
    size = op.arg;
  } else {
    is_free = true;
  }
}

// The pointer a published table entry v of `slot` holds: taken out of the
// table, tag checked; nullptr when the malloc failed.
__device__ inline void* take_entry(uint64_t* ptrs, uint32_t slot, uint64_t v, op_counts& c) {
  utils::atomic_store(&ptrs[slot], (uint64_t)0);
  if (v == kMallocFailed) {
    c.skipped += 1;
    return nullptr;
  }
  uint64_t* p = reinterpret_cast<uint64_t*>(v);
  c.bad += utils::atomic_load(p) != (uint64_t)slot;
  c.bytes += 16;  // tag written at malloc, read here
  c.pairs += 1;
  return p;
}

// The pointer a free op releases: taken out of the table (bounded wait
// for its malloc to publish it), tag checked; nullptr when the malloc
// failed or the wait ran out.
__device__ inline void* wait_free(uint64_t* ptrs, uint32_t slot, op_counts& c,
                                  uint32_t spin_cap, uint64_t* ctr) {
  uint64_t v = utils::atomic_load(&ptrs[slot]);
  for (uint32_t spins = 0; v == 0; ++spins) {  // its malloc has not run yet
    if (spins >= spin_cap) {
      utils::atomic_add(&ctr[GAL_CTR_TIMEOUTS], (uint64_t)1);
      return nullptr;
    }
    __builtin_amdgcn_s_sleep(4);
    v = utils::atomic_load(&ptrs[slot]);
    c.polls += 1;
  }
  return take_entry(ptrs, slot, v, c);
}

// Decode op i: the pointer to free (taken out of the table, tag checked) or
// the size to allocate.
__device__ inline void take_op(const gal_batch_op* ops, uint64_t i, bool active,
                               uint64_t* ptrs, uint64_t nslots, op_counts& c, void*& old,
                               bool& want, uint64_t& size, uint32_t& slot,
                               uint32_t spin_cap, uint64_t* ctr) {
  bool is_free = false;
  decode_op(ops, i, active, nslots, c, want, size, slot, is_free);
  old = is_free ? wait_free(ptrs, slot, c, spin_cap, ctr) : nullptr;
}

__device__ inline void give_op(void* np, bool want, uint32_t slot, uint64_t* ptrs,
                               op_counts& c) {
  if (want) {
    c.mallocs += 1;
    if (np == nullptr) c.misses += 1;
    if (np) utils::atomic_store(static_cast<uint64_t*>(np), (uint64_t)slot);
  }
  // the tag is acknowledged before the entry is published
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  if (want)
    utils::atomic_store(&ptrs[slot], np ? reinterpret_cast<uint64_t>(np) : kMallocFailed);
}

// One batch through the workgroup-batched path: a persistent grid walks
// chunks of kThreads ops.
template <int kThreads>
__global__ __launch_bounds__(kThreads) void k_batch_batched(gal_state* s, const gal_batch_op* ops,
                                                            uint64_t n, uint64_t* ptrs,
                                                            uint64_t nslots, uint64_t* ctr,
                                                            uint32_t spin_cap) {
  __shared__ batch::lds_state lds;
  batch::batch_init(lds);
  op_counts c;
  const uint64_t nchunks = (n + kThreads - 1) / kThreads;
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t i = chunk * kThreads + threadIdx.x;
    void* old;
    bool want;
    uint64_t size;
    uint32_t slot;
    take_op(ops, i, i < n, ptrs, nslots, c, old, want, size, slot, spin_cap, ctr);
    void* np = batch::batch_step(s, old, want, size, lds);
    give_op(np, want, slot, ptrs, c);
  }
  flush_counts(ctr, c);
}

// One batch with mallocs through the workgroup's LDS slice cache
// (slice_cache.hpp) and frees through the per-wave path.
template <int kThreads>
__global__ __launch_bounds__(kThreads) void k_batch_cached(gal_state* s, const gal_batch_op* ops,
                                                           uint64_t n, uint64_t* ptrs,
                                                           uint64_t nslots, uint64_t* ctr,
                                                           uint32_t spin_cap) {
  __shared__ cache::slice_cache lds;
  cache::cache_init(lds);
  op_counts c;
  const uint64_t nchunks = (n + kThreads - 1) / kThreads;
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t i = chunk * kThreads + threadIdx.x;
    void* old;
    bool want;
    uint64_t size;
    uint32_t slot;
    take_op(ops, i, i < n, ptrs, nslots, c, old, want, size, slot, spin_cap, ctr);
    if (old) detail::free_impl(s, old);
    void* np = cache::cache_malloc(s, lds, want, size);
    give_op(np, want, slot, ptrs, c);
  }
  cache::cache_flush(s, lds);
  flush_counts(ctr, c);
}

// The same batch through the per-thread path.
__global__ __launch_bounds__(256) void k_batch_plain(gal_state* s, const gal_batch_op* ops,
                                                     uint64_t n, uint64_t* ptrs, uint64_t nslots,
                                                     uint64_t* ctr, uint32_t spin_cap) {
  op_counts c;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t rounds = (n + stride - 1) / stride;
  for (uint64_t r = 0; r < rounds; ++r) {
    const uint64_t i = r * stride + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    void* old;
    bool want;
    uint64_t size;
    uint32_t slot;
    take_op(ops, i, i < n, ptrs, nslots, c, old, want, size, slot, spin_cap, ctr);
    if (old) detail::free_impl(s, old);
    void* np = nullptr;
    if (want) np = detail::malloc_impl(s, size);
    give_op(np, want, slot, ptrs, c);
  }
  flush_counts(ctr, c);
}

// Cleanup: free every pointer left in the table and clear every entry.
__device__ inline void take_left(uint64_t* ptrs, uint64_t i, bool in, op_counts& c,
                                 void*& old) {
  old = nullptr;
  if (!in) return;
  const uint64_t v = ptrs[i];
  if (v == 0) return;
  ptrs[i] = 0;
  if (v == kMallocFailed) return;
  uint64_t* p = reinterpret_cast<uint64_t*>(v);
  c.bad += *p != i;
  c.bytes += 16;
  c.pairs += 1;
  old = p;
}

template <int kThreads>
__global__ __launch_bounds__(kThreads) void k_table_free_batched(gal_state* s, uint64_t* ptrs,
                                                                 uint64_t nslots, uint64_t* ctr) {
  // Blocks the cleanup empties are returned together at the end (nothing
  // allocates meanwhile): returning each on the spot held the workgroup at
  // its next barrier for the one thread's ring return.
  constexpr uint32_t kLater = 4 * kThreads;
  __shared__ batch::lds_state lds;
  __shared__ uint32_t later[kLater];
  __shared__ uint32_t nlater;
  if (threadIdx.x == 0) nlater = 0;
  batch::batch_init(lds);
  op_counts c;
  const uint32_t my_p = detail::wave_part(s);
  const uint64_t nchunks = (nslots + kThreads - 1) / kThreads;
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t i = chunk * kThreads + threadIdx.x;
    void* old;
    take_left(ptrs, i, i < nslots, c, old);
    const bool grouped = batch::free_stage(s, old, lds, my_p);
    __syncthreads();
    batch::free_scan_later(s, lds, my_p, later, &nlater, kLater);
    const bool rest = old != nullptr && !grouped;
    if (__ballot(rest)) {
      if (rest) detail::free_impl(s, old);
    }
    __syncthreads();
  }
  batch::return_later(s, my_p, later, nlater, kLater);
  flush_counts(ctr, c);
}

__global__ __launch_bounds__(256) void k_table_free_plain(gal_state* s, uint64_t* ptrs,
                                                          uint64_t nslots, uint64_t* ctr) {
  op_counts c;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t rounds = (nslots + stride - 1) / stride;
  for (uint64_t r = 0; r < rounds; ++r) {
    const uint64_t i = r * stride + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    void* old;
    take_left(ptrs, i, i < nslots, c, old);
    if (old) detail::free_impl(s, old);
  }
  flush_counts(ctr, c);
}

// ---- one persistent launch for a whole stream ----
//
// The stream is cut into chunks of at most kChunkRounds * kThreads ops that
// never span two batches, so a free never waits on a malloc of its own
// chunk.  Workgroups take chunks in stream order from a queue counter and
// run them at once: a free waits only for its own table entry.  A chunk is
// only taken by a running workgroup, a free's malloc lies in a chunk taken
// earlier, and a workgroup serves a chunk's mallocs as soon as it takes the
// chunk, without waiting on any entry, so every entry is published and the
// grid always drains; the waits are bounded all the same (counter [5]).  A
// workgroup serves and publishes a chunk's mallocs before it runs the
// frees of its previous chunk, which shortens the chain of waits through
// the stream; the batched forms batch the mallocs across the workgroup and
// free per wave.
struct stream_chunk {
  uint64_t begin;  // first op
  uint32_t count;  // ops
  uint32_t batch;
};

constexpr uint32_t kChunkRounds = 4;

// kMode: 0 per-thread malloc, 1 workgroup-batched malloc (batched.hpp),
// 2 malloc through the workgroup's LDS slice cache (slice_cache.hpp).
template <int kThreads, int kMode>
__global__ __launch_bounds__(kThreads) void k_stream(gal_state* s, const gal_batch_op* ops,
                                                     const stream_chunk* chunks, uint32_t nchunks,
                                                     uint32_t* queue, uint64_t* ptrs,
                                                     uint64_t nslots, uint64_t* ctr,
                                                     uint32_t spin_cap, bool timed) {
  __shared__ batch::lds_state lds;
  __shared__ cache::slice_cache sc;
  __shared__ uint32_t ranks[kMode == 1 ? kChunkRounds * kThreads : 1];  // per op of a chunk
  __shared__ uint32_t cur;
  if (kMode == 1) batch::batch_init(lds);
  if (kMode == 2) cache::cache_init(sc);
  op_counts c;
  // A chunk's frees run after the next chunk's mallocs: they wait for
  // entries that other chunks publish, and meanwhile this workgroup serves
  // and publishes the entries that later chunks' frees wait for.
  stream_chunk prev{0, 0, 0};  // count 0: no frees pending
  for (;;) {
    if (threadIdx.x == 0) cur = utils::atomic_add(queue, 1u);
    __syncthreads();
    const uint32_t k = cur;
    __syncthreads();  // every thread has read cur before it is written again
    const bool done = k >= nchunks;
    const stream_chunk ch = done ? stream_chunk{0, 0, 0} : chunks[k];
    const uint32_t rounds = (ch.count + kThreads - 1) / kThreads;
    // Every malloc of the chunk is served and published before any free of
    // the workgroup waits: the frees of later chunks wait on exactly these
    // entries, and a chunk's frees only name entries of earlier batches.
    // Serving them round by round instead (malloc, then wait and free, per
    // round) held a chunk's later mallocs behind its own earlier waits, so
    // waits chained from chunk to chunk through the stream.
    const uint64_t t0 = timed ? wall_clock64() : 0;
    if (done) {
      // nothing to serve: only the last chunk's frees are left
    } else if constexpr (kMode == 1) {
      // The workgroup form counts the requests of all the chunk's rounds
      // first and claims them together: two barriers and one claim per
      // class for the chunk, where one batch_malloc per round took two
      // barriers and one claim per class each.
      const uint32_t p = detail::wave_part(s);
      for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t j = r * kThreads + threadIdx.x;
        bool want, is_free;
        uint64_t size;
        uint32_t slot;
        decode_op(ops, ch.begin + j, j < ch.count, nslots, c, want, size, slot, is_free);
        ranks[j] = batch::malloc_stage<true>(s, want, size, lds).rank;
      }
      __syncthreads();
      if (threadIdx.x < 64) batch::claim_classes(s, lds, p);
      __syncthreads();
      for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t j = r * kThreads + threadIdx.x;
        bool want = false;
        uint64_t size = 0;
        uint32_t slot = 0;
        if (j < ch.count) {  // decoded (and invalid ops counted) above
          const gal_batch_op op = ops[ch.begin + j];
          want = op.kind == 0 && op.slot < nslots;
          size = op.arg;
          slot = op.slot;
        }
        batch::malloc_req q = batch::malloc_classify<true>(s, want, size);
        q.rank = ranks[j];
        const uint64_t bytes = batch::malloc_take<true>(s, q, lds, p);
        void* np = bytes == kFail ? nullptr : (void*)(s->memory + bytes);
        const bool rest = want && bytes == kFail;
        if (__ballot(rest)) {
          void* w = nullptr;
          if (rest) w = detail::malloc_impl(s, size);
          if (rest) np = w;
        }
        give_op(np, want, slot, ptrs, c);
      }
    } else {
      for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t j = r * kThreads + threadIdx.x;
        bool want, is_free;
        uint64_t size;
        uint32_t slot;
        decode_op(ops, ch.begin + j, j < ch.count, nslots, c, want, size, slot, is_free);
        void* np = nullptr;
        if (kMode == 2)
          np = cache::cache_malloc(s, sc, want, size);
        else if (want)
          np = detail::malloc_impl(s, size);
        give_op(np, want, slot, ptrs, c);
      }
    }
    if (timed) c.malloc_ticks += wall_clock64() - t0;
    // Frees go through the wave path in every form: a workgroup-wide free
    // would hold every wave behind the slowest wait.  (Setting aside frees
    // whose entry is not yet published, to wait for them after the chunk's
    // other frees, moved wave time from waiting to freeing and did not
    // shorten a replay.)
    const uint32_t prev_rounds = (prev.count + kThreads - 1) / kThreads;
    for (uint32_t r = 0; r < prev_rounds; ++r) {
      const uint32_t j = r * kThreads + threadIdx.x;
      bool is_free = false;
      uint32_t slot = 0;
      if (j < prev.count) {
        const gal_batch_op op = ops[prev.begin + j];
        is_free = op.kind == 1 && op.slot < nslots;  // invalid ops were counted above
        slot = op.slot;
      }
      const uint64_t t1 = timed ? wall_clock64() : 0;
      void* old = is_free ? wait_free(ptrs, slot, c, spin_cap, ctr) : nullptr;
      const uint64_t t2 = timed ? wall_clock64() : 0;
      if (old) detail::free_impl(s, old);
      if (timed) {
        c.wait_ticks += t2 - t1;
        c.free_ticks += wall_clock64() - t2;
      }
    }
    if (done) break;
    prev = ch;
  }
  if (kMode == 2) cache::cache_flush(s, sc);
  flush_counts(ctr, c, timed);
}

__global__ void k_stream_reset(uint32_t* queue) {
  if (blockIdx.x == 0 && threadIdx.x == 0) queue[0] = 0;
}

// The chunk table of a stream from its batch offsets, the same table
// gal_batch_plan_create builds on the host: the ops of batch b cut into
// chunks of at most `per`, in stream order.  One workgroup of 1024 threads
// takes 1024 batches at a time, one per thread, and places their chunks by
// an LDS prefix sum of the chunk counts.  Also resets the queue counter.
__global__ __launch_bounds__(1024) void k_build_chunks(const uint64_t* offsets, uint32_t nbatches,
                                                      uint32_t per, stream_chunk* chunks,
                                                      uint32_t* queue) {
  __shared__ uint32_t part[1024];
  __shared__ uint32_t base_s;
  if (threadIdx.x == 0) {
    base_s = 0;
    queue[0] = 0;
  }
  __syncthreads();
  for (uint32_t b0 = 0; b0 < nbatches; b0 += 1024) {
    const uint32_t b = b0 + threadIdx.x;
    uint64_t lo = 0, n = 0;
    if (b < nbatches) {
      lo = offsets[b];
      n = offsets[b + 1] - lo;
    }
    const uint32_t c = (uint32_t)((n + per - 1) / per);
    part[threadIdx.x] = c;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {  // inclusive scan
      const uint32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
      __syncthreads();
      part[threadIdx.x] += v;
      __syncthreads();
    }
    const uint32_t at = base_s + part[threadIdx.x] - c;
    for (uint32_t k = 0; k < c; ++k) {
      const uint64_t left = n - (uint64_t)k * per;
      chunks[at + k] = stream_chunk{lo + (uint64_t)k * per, (uint32_t)(left < per ? left : per), b};
    }
    __syncthreads();
    if (threadIdx.x == 1023) base_s += part[1023];
    __syncthreads();
  }
}

// Workspace of gal_run_stream_static, the entry point that allocates
// nothing: the queue counter, a copy of the batch offsets and the chunk
// table, in one array per device that lives until the process ends.  Host
// code reaches it only through the address hipGetSymbolAddress returns for
// the current device (the array's name is a host-side stand-in there).
constexpr uint32_t kStaticMaxBatches = 65535;
constexpr uint32_t kStaticMaxChunks = 1u << 17;
constexpr size_t kScratchOffsets = 256;  // bytes [0, 256): the queue counter
constexpr size_t kScratchChunks = kScratchOffsets + 8 * (size_t)(kStaticMaxBatches + 1);
constexpr size_t kScratchBytes = kScratchChunks + sizeof(stream_chunk) * (size_t)kStaticMaxChunks;
static_assert(sizeof(stream_chunk) == 16, "chunk table entries are 16 bytes");
static __device__ __align__(256) uint64_t scratch[kScratchBytes / 8];

template <typename K>
unsigned fill_grid(K kernel, int threads, uint32_t workgroups) {
  if (workgroups) return workgroups;
  int dev = 0, cus = 256, per_cu = 1;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return (unsigned)(cus * per_cu);
}

constexpr uint32_t kStreamSpinCap = 1u << 22;  // polls of a free waiting for its entry

template <int kThreads>
void launch_batched(gal_state* s, const gal_batch_op* ops, uint64_t n, uint64_t* ptrs,
                    uint64_t nslots, uint64_t* ctr, uint32_t workgroups, hipStream_t st) {
  unsigned g = fill_grid(k_batch_batched<kThreads>, kThreads, workgroups);
  const uint64_t chunks = (n + kThreads - 1) / kThreads;
  if (g > chunks) g = (unsigned)chunks;
  k_batch_batched<kThreads><<<g, kThreads, 0, st>>>(s, ops, n, ptrs, nslots, ctr, kStreamSpinCap);
}

template <int kThreads>
void launch_cached(gal_state* s, const gal_batch_op* ops, uint64_t n, uint64_t* ptrs,
                   uint64_t nslots, uint64_t* ctr, uint32_t workgroups, hipStream_t st) {
  unsigned g = fill_grid(k_batch_cached<kThreads>, kThreads, workgroups);
  const uint64_t chunks = (n + kThreads - 1) / kThreads;
  if (g > chunks) g = (unsigned)chunks;
  k_batch_cached<kThreads><<<g, kThreads, 0, st>>>(s, ops, n, ptrs, nslots, ctr, kStreamSpinCap);
}

template <int kThreads>
void launch_cleanup(gal_state* s, uint64_t* ptrs, uint64_t nslots, uint64_t* ctr,
                    uint32_t workgroups, hipStream_t st) {
  unsigned g = fill_grid(k_table_free_batched<kThreads>, kThreads, workgroups);
  const uint64_t chunks = (nslots + kThreads - 1) / kThreads;
  if (g > chunks) g = (unsigned)chunks;
  k_table_free_batched<kThreads><<<g, kThreads, 0, st>>>(s, ptrs, nslots, ctr);
}

}  // namespace

struct gal_batch_plan {
  gal_allocator* a = nullptr;
  hipStream_t capture = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // persistent-launch state (device)
  stream_chunk* chunks = nullptr;
  uint32_t* queue = nullptr;
};

// Enqueue reset, the persistent launch and the cleanup on `st`.
static int enqueue_persistent(gal_state* s, const stream_chunk* chunks, uint32_t* queue,
                              const gal_batch_op* ops, uint32_t nchunks, uint64_t* ptrs,
                              uint64_t nslots, uint64_t* ctr, uint32_t threads, uint32_t batched,
                              uint32_t workgroups, hipStream_t st) {
  const bool timed = (batched & GAL_BATCH_TIMED) != 0;
  batched &= ~GAL_BATCH_TIMED;
  k_stream_reset<<<1, 64, 0, st>>>(queue);
  if (nchunks) {
#define GAL_STREAM_LAUNCH(T, B)                                                              \
  do {                                                                                        \
    unsigned g = fill_grid(k_stream<T, B>, T, workgroups);                                    \
    if (g > nchunks) g = nchunks;                                                             \
    k_stream<T, B><<<g, T, 0, st>>>(s, ops, chunks, nchunks, queue, ptrs, nslots, ctr,       \
                                    kStreamSpinCap, timed);                                   \
  } while (0)
    if (batched == 0)
      GAL_STREAM_LAUNCH(256, 0);
    else if (batched == 1 && threads == 256)
      GAL_STREAM_LAUNCH(256, 1);
    else if (batched == 1 && threads == 512)
      GAL_STREAM_LAUNCH(512, 1);
    else if (batched == 1)
      GAL_STREAM_LAUNCH(1024, 1);
    else if (threads == 256)
      GAL_STREAM_LAUNCH(256, 2);
    else if (threads == 512)
      GAL_STREAM_LAUNCH(512, 2);
    else
      GAL_STREAM_LAUNCH(1024, 2);
#undef GAL_STREAM_LAUNCH
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "stream launch");
  if (nslots) {
    if (batched != 1) {
      unsigned g = fill_grid(k_table_free_plain, 256, workgroups);
      const uint64_t need = (nslots + 255) / 256;
      if (g > need) g = (unsigned)need;
      k_table_free_plain<<<g, 256, 0, st>>>(s, ptrs, nslots, ctr);
    } else if (threads == 256) {
      launch_cleanup<256>(s, ptrs, nslots, ctr, workgroups, st);
    } else if (threads == 512) {
      launch_cleanup<512>(s, ptrs, nslots, ctr, workgroups, st);
    } else {
      launch_cleanup<1024>(s, ptrs, nslots, ctr, workgroups, st);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "cleanup launch");
  }
  return GAL_OK;
}

// Enqueue every batch and the cleanup on `st`.
static int enqueue_stream(gal_allocator* a, const gal_batch_op* ops, const uint64_t* offsets,
                          uint32_t nbatches, uint64_t* ptrs, uint64_t nslots, uint64_t* ctr,
                          uint32_t threads, uint32_t batched, uint32_t workgroups,
                          hipStream_t st) {
  gal_state* s = a->dev;
  for (uint32_t b = 0; b < nbatches; ++b) {
    const uint64_t n = offsets[b + 1] - offsets[b];
    if (n == 0) continue;
    const gal_batch_op* o = ops + offsets[b];
    if (batched == 0) {
      unsigned g = fill_grid(k_batch_plain, 256, workgroups);
      const uint64_t need = (n + 255) / 256;
      if (g > need) g = (unsigned)need;
      k_batch_plain<<<g, 256, 0, st>>>(s, o, n, ptrs, nslots, ctr, kStreamSpinCap);
    } else if (batched == 1 && threads == 256) {
      launch_batched<256>(s, o, n, ptrs, nslots, ctr, workgroups, st);
    } else if (batched == 1 && threads == 512) {
      launch_batched<512>(s, o, n, ptrs, nslots, ctr, workgroups, st);
    } else if (batched == 1) {
      launch_batched<1024>(s, o, n, ptrs, nslots, ctr, workgroups, st);
    } else if (threads == 256) {
      launch_cached<256>(s, o, n, ptrs, nslots, ctr, workgroups, st);
    } else if (threads == 512) {
      launch_cached<512>(s, o, n, ptrs, nslots, ctr, workgroups, st);
    } else {
      launch_cached<1024>(s, o, n, ptrs, nslots, ctr, workgroups, st);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "batch launch");
  }
  if (nslots) {
    if (batched != 1) {
      unsigned g = fill_grid(k_table_free_plain, 256, workgroups);
      const uint64_t need = (nslots + 255) / 256;
      if (g > need) g = (unsigned)need;
      k_table_free_plain<<<g, 256, 0, st>>>(s, ptrs, nslots, ctr);
    } else if (threads == 256) {
      launch_cleanup<256>(s, ptrs, nslots, ctr, workgroups, st);
    } else if (threads == 512) {
      launch_cleanup<512>(s, ptrs, nslots, ctr, workgroups, st);
    } else {
      launch_cleanup<1024>(s, ptrs, nslots, ctr, workgroups, st);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "cleanup launch");
  }
  return GAL_OK;
}

// The address of `scratch` on the current device, looked up once per
// device with that device current and kept.  Every call on a device shares
// the one array.
static std::mutex g_workspace_mu;
static char* g_workspace[64];

static int static_workspace(char** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= 64) return fail(GAL_ERR_INVALID, "device index beyond 63");
  std::lock_guard<std::mutex> lk(g_workspace_mu);
  if (!g_workspace[dev]) {
    char* ptr = nullptr;
    e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_workspace[dev] = ptr;
  }
  *out = g_workspace[dev];
  return GAL_OK;
}

static int check_plan_args(gal_allocator* a, const gal_batch_op* ops, const uint64_t* offsets,
                           uint32_t nbatches, uint64_t* ptrs, uint64_t nslots, uint64_t* ctr,
                           uint32_t* threads, uint32_t batched) {
  if (!a || !offsets || !ctr || (nslots && !ptrs)) return fail(GAL_ERR_INVALID, "null argument");
  if ((batched & ~GAL_BATCH_TIMED) > 2)
    return fail(GAL_ERR_INVALID, "batched is 0 (per thread), 1 (workgroup batches) or 2 (slice "
                                 "cache), with GAL_BATCH_TIMED or-ed in for the timed kernel");
  if (nbatches && offsets[nbatches] && !ops) return fail(GAL_ERR_INVALID, "null ops");
  for (uint32_t b = 0; b < nbatches; ++b)
    if (offsets[b + 1] < offsets[b]) return fail(GAL_ERR_INVALID, "offsets must not decrease");
  if (nslots > 0xFFFFFFFFull) return fail(GAL_ERR_INVALID, "at most 2^32 slots");
  if (*threads == 0) *threads = 1024;
  if (*threads != 256 && *threads != 512 && *threads != 1024)
    return fail(GAL_ERR_INVALID, "batched streams run 256, 512 or 1024 threads per workgroup");
  return GAL_OK;
}

extern "C" {

int gal_check_batch_stream(const gal_batch_op* ops, const uint64_t* offsets, uint32_t nbatches,
                           uint64_t nslots, int persistent) {
  if (!offsets || (nbatches && offsets[nbatches] && !ops))
    return fail(GAL_ERR_INVALID, "null argument");
  if (nslots > 0xFFFFFFFFull) return fail(GAL_ERR_INVALID, "at most 2^32 slots");
  for (uint32_t b = 0; b < nbatches; ++b)
    if (offsets[b + 1] < offsets[b]) return fail(GAL_ERR_INVALID, "offsets must not decrease");
  // per slot: batch of its last malloc (or ~0), and whether that malloc's
  // pointer is still in the table (freed or not)
  constexpr uint32_t kNone = 0xFFFFFFFFu;
  std::vector<uint32_t> last(nslots, kNone);
  std::vector<uint8_t> live(nslots, 0), used(nslots, 0);
  for (uint32_t b = 0; b < nbatches; ++b) {
    for (uint64_t i = offsets[b]; i < offsets[b + 1]; ++i) {
      const gal_batch_op& op = ops[i];
      if (op.kind > 1) return fail(GAL_ERR_INVALID, "op " + std::to_string(i) + ": unknown kind");
      if (op.slot >= nslots)
        return fail(GAL_ERR_INVALID, "op " + std::to_string(i) + ": slot beyond nslots");
      const uint32_t sl = op.slot;
      if (op.kind == 0) {
        if (persistent && used[sl])
          return fail(GAL_ERR_INVALID, "op " + std::to_string(i) + ": slot " +
                                           std::to_string(sl) +
                                           " is filled a second time (a persistent stream "
                                           "fills each slot at most once)");
        if (live[sl] || last[sl] == b)
          return fail(GAL_ERR_INVALID, "op " + std::to_string(i) + ": slot " +
                                           std::to_string(sl) +
                                           " is filled again before its pointer was freed");
        last[sl] = b;
        live[sl] = 1;
        used[sl] = 1;
      } else {
        if (!live[sl] || last[sl] >= b)
          return fail(GAL_ERR_INVALID, "op " + std::to_string(i) + ": slot " +
                                           std::to_string(sl) +
                                           " is freed without a malloc in an earlier batch");
        live[sl] = 0;
      }
    }
  }
  return GAL_OK;
}

int gal_run_batches(gal_allocator* a, const gal_batch_op* ops, const uint64_t* offsets,
                    uint32_t nbatches, uint64_t* ptrs, uint64_t nslots, uint64_t* ctr,
                    uint32_t threads, uint32_t batched, uint32_t workgroups, void* stream) {
  const int rc = check_plan_args(a, ops, offsets, nbatches, ptrs, nslots, ctr, &threads, batched);
  if (rc != GAL_OK) return rc;
  // one launch per batch: GAL_BATCH_TIMED times only the persistent kernel
  return enqueue_stream(a, ops, offsets, nbatches, ptrs, nslots, ctr, threads,
                        batched & ~GAL_BATCH_TIMED, workgroups, (hipStream_t)stream);
}

int gal_batch_plan_create(gal_allocator* a, const gal_batch_op* ops, const uint64_t* offsets,
                          uint32_t nbatches, uint64_t* ptrs, uint64_t nslots, uint64_t* ctr,
                          uint32_t threads, uint32_t batched, uint32_t workgroups,
                          gal_batch_plan** out) {
  if (!out) return fail(GAL_ERR_INVALID, "null argument");
  *out = nullptr;
  int rc = check_plan_args(a, ops, offsets, nbatches, ptrs, nslots, ctr, &threads, batched);
  if (rc != GAL_OK) return rc;
  // The persistent launch runs a stream's chunks at once: a slot filled
  // twice could be overwritten before its first pointer is freed, and a
  // free whose malloc is not in an earlier batch would wait in vain.  The
  // ops are checked on the host once, when the plan is made.
  if (nbatches && offsets[nbatches] > offsets[0]) {
    std::vector<gal_batch_op> host(offsets[nbatches]);
    const hipError_t ce = hipMemcpy(host.data(), ops, host.size() * sizeof(gal_batch_op),
                                    hipMemcpyDeviceToHost);
    if (ce != hipSuccess) return hip_fail(ce, "copy of the ops for checking");
    rc = gal_check_batch_stream(host.data(), offsets, nbatches, nslots, 1);
    if (rc != GAL_OK) return rc;
  }
  // chunk table: at most kChunkRounds * threads ops, never across batches
  const uint32_t per = kChunkRounds * ((batched & ~GAL_BATCH_TIMED) ? threads : 256u);
  std::vector<stream_chunk> chunks;
  for (uint32_t b = 0; b < nbatches; ++b) {
    for (uint64_t at = offsets[b]; at < offsets[b + 1]; at += per) {
      const uint64_t left = offsets[b + 1] - at;
      chunks.push_back({at, (uint32_t)(left < per ? left : per), b});
    }
  }
  if (chunks.size() > 0xFFFFFFF0ull) return fail(GAL_ERR_INVALID, "stream too long");
  gal_batch_plan* plan = new gal_batch_plan();
  plan->a = a;
  hipError_t e = hipSuccess;
  if ((e = hipMalloc(&plan->chunks, (chunks.size() + 1) * sizeof(stream_chunk))) != hipSuccess ||
      (e = hipMalloc(&plan->queue, 4 * sizeof(uint32_t))) != hipSuccess ||
      (!chunks.empty() &&
       (e = hipMemcpy(plan->chunks, chunks.data(), chunks.size() * sizeof(stream_chunk),
                      hipMemcpyHostToDevice)) != hipSuccess)) {
    gal_batch_plan_destroy(plan);
    return hip_fail(e, "plan buffers");
  }
  e = hipStreamCreateWithFlags(&plan->capture, hipStreamNonBlocking);
  if (e != hipSuccess) {
    gal_batch_plan_destroy(plan);
    return hip_fail(e, "capture stream");
  }
  e = hipStreamBeginCapture(plan->capture, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    gal_batch_plan_destroy(plan);
    return hip_fail(e, "begin capture");
  }
  rc = enqueue_persistent(a->dev, plan->chunks, plan->queue, ops, (uint32_t)chunks.size(), ptrs,
                          nslots, ctr, threads, batched, workgroups, plan->capture);
  e = hipStreamEndCapture(plan->capture, &plan->graph);
  if (rc == GAL_OK && e != hipSuccess) rc = hip_fail(e, "end capture");
  if (rc == GAL_OK) {
    e = hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) rc = hip_fail(e, "graph instantiate");
  }
  if (rc != GAL_OK) {
    gal_batch_plan_destroy(plan);
    return rc;
  }
  *out = plan;
  return GAL_OK;
}

int gal_run_stream_static(gal_allocator* a, const gal_batch_op* ops, const uint64_t* offsets,
                          uint32_t nbatches, uint64_t* ptrs, uint64_t nslots, uint64_t* ctr,
                          uint32_t threads, uint32_t batched, uint32_t workgroups, void* stream) {
  int rc = check_plan_args(a, ops, offsets, nbatches, ptrs, nslots, ctr, &threads, batched);
  if (rc != GAL_OK) return rc;
  if (nbatches > kStaticMaxBatches)
    return fail(GAL_ERR_INVALID, "the static workspace holds at most 65535 batches; got " +
                                     std::to_string(nbatches));
  const uint32_t per = kChunkRounds * ((batched & ~GAL_BATCH_TIMED) ? threads : 256u);
  uint64_t nchunks = 0;
  for (uint32_t b = 0; b < nbatches; ++b) nchunks += (offsets[b + 1] - offsets[b] + per - 1) / per;
  if (nchunks > kStaticMaxChunks)
    return fail(GAL_ERR_INVALID, "the static workspace holds at most 131072 chunks of " +
                                     std::to_string(per) + " ops; the stream needs " +
                                     std::to_string(nchunks));
  char* sp = nullptr;
  rc = static_workspace(&sp);
  if (rc != GAL_OK) return rc;
  uint32_t* queue = reinterpret_cast<uint32_t*>(sp);
  uint64_t* offs = reinterpret_cast<uint64_t*>(sp + kScratchOffsets);
  stream_chunk* chunks = reinterpret_cast<stream_chunk*>(sp + kScratchChunks);
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipMemcpyAsync(offs, offsets, 8 * ((size_t)nbatches + 1), hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return hip_fail(e, "offsets into the static workspace");
  if (nbatches) k_build_chunks<<<1, 1024, 0, st>>>(offs, nbatches, per, chunks, queue);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "chunk table launch");
  return enqueue_persistent(a->dev, chunks, queue, ops, (uint32_t)nchunks, ptrs, nslots, ctr,
                            threads, batched, workgroups, st);
}

int gal_batch_plan_launch(gal_batch_plan* plan, void* stream) {
  if (!plan || !plan->exec) return fail(GAL_ERR_INVALID, "null plan");
  const hipError_t e = hipGraphLaunch(plan->exec, (hipStream_t)stream);
  return e == hipSuccess ? GAL_OK : hip_fail(e, "graph launch");
}

int gal_batch_plan_destroy(gal_batch_plan* plan) {
  if (!plan) return GAL_OK;
  if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
  if (plan->graph) (void)hipGraphDestroy(plan->graph);
  if (plan->capture) (void)hipStreamDestroy(plan->capture);
  if (plan->chunks) (void)hipFree(plan->chunks);
  if (plan->queue) (void)hipFree(plan->queue);
  delete plan;
  return GAL_OK;
}

}  // extern "C"

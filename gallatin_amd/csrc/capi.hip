// Host lifecycle of device allocators behind the C ABI of gallatin_capi.h:
// geometry checks, one metadata arena + one pool per instance, init and boot
// kernels, drain, snapshots and statistics.
//
// Reference counterparts: Gallatin::generate_on_device / free_on_device /
// print_info (gallatin.cuh:324-421, 643-668, 1855-2011), veb_tree and
// alloc_table construction (veb.cuh:595-649, memory_table.cuh:471-647), the
// boot kernels (gallatin.cuh:136-221) and calculate_overhead
// (gallatin.cuh:1912-1944).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "gallatin/allocators/gallatin.hpp"
#include "gallatin_capi.h"
#include "internal.hpp"

using namespace gallatin;
using namespace gallatin::allocators;

namespace gallatin_capi {

static thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(GAL_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

static std::mutex g_registry_mu;
static std::map<const void*, gal_allocator*> g_registry;

}  // namespace gallatin_capi

using namespace gallatin_capi;

// ------------------------------------------------------------ init kernels

__global__ void k_init_segments(gal_state* s) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s->num_segments) return;
  s->seg_state[i] = kSegFreeState;  // gallatin_init_counters_kernel
  s->seg_queue[i] = gal_seg_queue{0, 0, 0};
}

__global__ void k_init_blocks(gal_state* s) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (s->num_segments << s->bps_shift)) return;
  *detail::block_at(s, (uint32_t)i) = gal_block{4097u, 0u};  // Block::init, block.cuh:64-68
  s->ring[i] = kNullBlock;
}

// init_bits (veb.cuh:72-84) for every layer of the segment tree: bits below
// the layer universe set.
__global__ void k_fill_tree(gal_tree t) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t u = t.universe;
  for (int l = (int)t.num_layers - 1; l >= 0; --l) {
    const uint64_t nw = t.layer_words[l];
    if (i < nw) {
      uint64_t v = ~0ULL;
      if (i == nw - 1 && (u & 63)) v = (1ULL << (u & 63)) - 1;
      t.words[t.layer_off[l] + i] = v;
    }
    u = (u - 1) / 64 + 1;
  }
}

// boot_shared_block_container_one_thread order (gallatin.cuh:195-221), run
// by one wave so tree searches are cooperative.  Slot i of a class belongs
// to partition i / (slots / partitions) and takes its block from a segment
// homed there (the kernel boundary writes every L2 back before use).
__global__ void k_boot(gal_state* s) {
  for (uint32_t t = 0; t < s->num_trees; ++t) {
    const uint32_t per = s->slot_count[t] / s->num_parts;
    for (uint32_t i = 0; i < s->slot_count[t]; ++i) {
      const uint32_t blk = detail::request_new_block(s, t, i / per, true, s->lim_request);
      if (blk == kNullBlock) {
        if (wave::is_first()) utils::atomic_store(&s->stats[kStatBootFailed], 1ULL);
        return;
      }
      if (wave::is_first()) utils::atomic_store(&s->slots[s->slot_off[t] + i], blk);
    }
  }
}

// Quiescent drain: pinned blocks give back their unclaimed slices, then
// formatted segments with every block free are released as return_block
// would release them.  Runs after a kernel boundary, so every L2 has been
// written back and partitions play no part.
__global__ void k_drain(gal_state* s) {
  for (uint32_t t = 0; t < s->num_trees; ++t) {
    for (uint32_t i = 0; i < s->slot_count[t]; ++i) {
      if (wave::is_first()) {
        const uint32_t blk = utils::atomic_exch(&s->slots[s->slot_off[t] + i], kNullBlock);
        if (blk < kRefilling) {
          uint32_t c = utils::atomic_load(&detail::block_at(s, blk)->malloc_counter) & kCountMask;
          if (c > 4096u) c = 4096u;
          // act as the block's home so a completed segment is released
          const uint64_t st = utils::atomic_load(&s->seg_state[blk >> s->bps_shift]);
          if (c < 4096u) detail::block_release(s, blk, 4096u - c, seg_home(st));
        }
      }
    }
  }
  const uint64_t act = wave::active();
  const uint32_t k = wave::count(act);
  const uint32_t rank = wave::count(act & wave::lanes_below());
  for (uint64_t base = 0; base < s->num_segments; base += k) {
    const uint64_t seg = base + rank;
    bool cand = false;
    uint64_t st = 0;
    if (seg < s->num_segments) {
      st = utils::atomic_load(&s->seg_state[seg]);
      const uint32_t t = (uint32_t)(st & 0xFFFF);
      if (t < s->num_trees) {
        const int64_t nb = (int64_t)(s->blocks_per_segment >> t);
        cand = (int64_t)(int32_t)(st >> 32) == nb - 1;
      }
    }
    for (uint64_t m = __ballot(cand); m; m &= m - 1) {
      const int l = wave::first(m);
      const uint64_t sg = wave::readlane(seg, l);
      const uint64_t sv = wave::readlane(st, l);
      if (wave::is_first()) {
        const uint32_t t = (uint32_t)(sv & 0xFFFF);
        const uint64_t closed = (0xFFFFFFFFULL << 32) | (sv & kSegIdMask);
        if (utils::atomic_cas(&s->seg_state[sg], sv, closed) == sv) {
          veb::remove(detail::class_tree(s, t, seg_home(sv)), sg);
          detail::clear_chunk_id(s, sg);
          veb::insert(s->seg_tree, sg);
          detail::add_stat(s, kStatSegmentsReleased, 1);
        }
      }
    }
  }
}

// --------------------------------------------------------------- helpers

static bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }
static uint32_t log2u(uint64_t x) { return 63u - (uint32_t)__builtin_clzll(x); }
static uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// Tree sizing (veb.cuh:595-618).
static void size_tree(gal_tree& t, uint64_t universe, uint64_t& word_cursor) {
  std::memset(&t, 0, sizeof(t));
  const int max_height = 64 - __builtin_clzll(universe);
  t.num_layers = (uint32_t)((max_height - 1) / 6 + 1);
  t.universe = universe;
  uint64_t u = universe;
  uint64_t words[kMaxLayers];
  for (uint32_t i = 0; i < t.num_layers; ++i) {
    words[t.num_layers - 1 - i] = (u - 1) / 64 + 1;
    u = (u - 1) / 64 + 1;
  }
  uint64_t off = 0;
  for (uint32_t l = 0; l < t.num_layers; ++l) {
    t.layer_off[l] = off;
    t.layer_words[l] = words[l];
    off += words[l];
  }
  t.words = reinterpret_cast<uint64_t*>(word_cursor);  // offset for now
  word_cursor += off;
}

static uint64_t tree_total_words(const gal_tree& t) {
  uint64_t n = 0;
  for (uint32_t l = 0; l < t.num_layers; ++l) n += t.layer_words[l];
  return n;
}

// ------------------------------------------------------------------ C ABI

extern "C" {

const char* gal_last_error(void) { return g_last_error.c_str(); }

void gal_config_default(gal_config* cfg, uint64_t max_bytes, uint64_t seed) {
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->max_bytes = max_bytes;
  cfg->bytes_per_segment = 16ULL << 20;
  cfg->smallest = 16;
  cfg->biggest = 4096;
  cfg->seed = seed;
  cfg->flags = 0;
  cfg->slots0 = 128;   // blocks_per_pinned_block, gallatin.cuh:346
  cfg->slots_min = 4;  // MIN_PINNED_CUTOFF, gallatin.cuh:124
  cfg->print_info = 0;
}

// The pool in device memory, pinned mapped host memory or managed memory.
static hipError_t alloc_pool(gal_allocator* a) {
  if (a->pool_kind != 0) {
    int dev = 0, ok = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess)
      e = hipDeviceGetAttribute(&ok,
                                a->pool_kind == 1 ? hipDeviceAttributeCanMapHostMemory
                                                  : hipDeviceAttributeManagedMemory,
                                dev);
    if (e != hipSuccess) return e;
    if (!ok) return hipErrorNotSupported;
  }
  if (a->pool_kind == 1) {
    void* host = nullptr;
    hipError_t e = hipHostMalloc(&host, a->pool_bytes, hipHostMallocMapped);
    if (e != hipSuccess) return e;
    void* dev = nullptr;
    e = hipHostGetDevicePointer(&dev, host, 0);
    if (e != hipSuccess || dev != host) {  // one address for host and device
      (void)hipHostFree(host);
      return e != hipSuccess ? e : hipErrorInvalidValue;
    }
    a->pool = host;
    return hipSuccess;
  }
  if (a->pool_kind == 2) return hipMallocManaged(&a->pool, a->pool_bytes);
  return hipMalloc(&a->pool, a->pool_bytes);
}

static hipError_t free_pool(gal_allocator* a) {
  return a->pool_kind == 1 ? hipHostFree(a->pool) : hipFree(a->pool);
}

int gal_create(const gal_config* cfg, gal_allocator** out) {
  if (!cfg || !out) return fail(GAL_ERR_INVALID, "null argument");
  *out = nullptr;
  if ((cfg->flags & GAL_FLAG_LAZY_BOOT) && (cfg->flags & GAL_FLAG_DETERMINISTIC))
    return fail(GAL_ERR_INVALID, "the deterministic mode boots its pinned blocks eagerly");
  if ((cfg->flags & GAL_FLAG_POOL_HOST) && (cfg->flags & GAL_FLAG_POOL_MANAGED))
    return fail(GAL_ERR_INVALID, "a pool is either pinned host or managed memory, not both");
  const uint64_t S = cfg->bytes_per_segment, m = cfg->smallest, M = cfg->biggest;
  if (!is_pow2(S) || !is_pow2(m) || !is_pow2(M) || m > M)
    return fail(GAL_ERR_INVALID, "segment, smallest and biggest must be powers of two, smallest <= biggest");
  if (S < M * kSlicesPerBlock)
    return fail(GAL_ERR_INVALID, "bytes_per_segment must be >= biggest * 4096 (gallatin.cuh:300)");
  const uint64_t N = cfg->max_bytes / S;
  const uint32_t T = log2u(M) - log2u(m) + 1;
  if (N == 0) return fail(GAL_ERR_INVALID, "pool smaller than one segment");
  if (T > (uint32_t)kMaxTrees) return fail(GAL_ERR_INVALID, "too many size classes");
  if (N >= (1ULL << 24) || N + T + 2 >= kFreeTree)
    return fail(GAL_ERR_INVALID, "too many segments");
  const uint64_t B = S / (m * kSlicesPerBlock);
  if (N * B >= 0xFFFFFFFFULL) return fail(GAL_ERR_INVALID, "too many blocks");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(GAL_ERR_NO_DEVICE, "no HIP device");

  gal_allocator* a = new gal_allocator();
  std::memset(&a->host, 0, sizeof(a->host));
  (void)hipGetDevice(&a->device);
  gal_state& h = a->host;
  h.bytes_per_segment = S;
  h.smallest = m;
  h.biggest = M;
  h.num_segments = N;
  h.blocks_per_segment = B;
  h.seed = cfg->seed;
  h.seg_shift = log2u(S);
  h.smallest_bits = log2u(m);
  h.bps_shift = log2u(B);
  h.num_trees = T;
  h.block_tree_bits = log2u(m * kSlicesPerBlock);
  h.flags = cfg->flags;
  h.lim_request = cfg->request_attempts ? cfg->request_attempts : (uint32_t)kRequestBlockAttempts;
  h.lim_slice = cfg->slice_attempts ? cfg->slice_attempts : (uint32_t)kSliceAttempts;
  h.lim_block = cfg->block_attempts ? cfg->block_attempts : (uint32_t)kBlockAttempts;
  h.lim_segment = cfg->segment_attempts ? cfg->segment_attempts : (uint32_t)kSegmentAttempts;
  // XCD partitions: one per XCD (32 CUs each on MI355X) or none.  A free
  // into a segment of its own partition only releases its payload to its
  // XCD's L2, which is sound only when a partition is exactly one XCD: with
  // two XCDs in a partition, a slice freed on one could be reused on the
  // other while the first XCD's L2 still holds the old payload as dirty
  // lines.  So P is the XCD count or 1.  Booting pins blocks of every class
  // in every partition; a pool where that would take more than about a
  // quarter of its segments gets one partition.
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, a->device) != hipSuccess)
    cus = 32;
  const uint32_t xcds = cus >= 64 ? (uint32_t)(cus / 32) : 1u;
  uint32_t P = cfg->partitions;
  if (cfg->flags & kFlagDeterministic) {
    P = 1;
  } else if (P == 0) {
    P = (uint64_t)xcds <= N / (4ULL * T) ? xcds : 1u;
  } else if (P != 1 && P != xcds) {
    delete a;
    return fail(GAL_ERR_INVALID, "partitions must be 0 (automatic), 1 or the number of XCDs");
  }
  if (P < 1 || P > (uint32_t)kMaxParts) P = 1;  // a partition per XCD, or one
  h.num_parts = P;
  // pinned slot counts, shared_block_storage.cuh:251-274, each rounded up to
  // a multiple of the partitions
  uint32_t p = cfg->slots0 ? cfg->slots0 : 1, total = 0;
  const uint32_t pmin = cfg->slots_min ? cfg->slots_min : 1;
  for (uint32_t t = 0; t < T; ++t) {
    const uint32_t c = (p + P - 1) / P * P;
    h.slot_off[t] = total;
    h.slot_count[t] = c;
    total += c;
    p /= 2;
    if (p < pmin) p = pmin;
  }
  h.slot_off[T] = total;
  h.total_slots = total;
  // trees: segment tree + one per class and partition, in one word arena
  uint64_t words = 0;
  size_tree(h.seg_tree, N, words);
  for (uint32_t t = 0; t < T * P; ++t) size_tree(h.sub_trees[t], N, words);
  a->tree_words = words;
  // metadata arena
  uint64_t off = 0;
  auto take = [&](uint64_t bytes) {
    const uint64_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  const uint64_t o_state = take(sizeof(gal_state));
  const uint64_t o_seg = take(N * sizeof(uint64_t));
  const uint64_t o_queue = take(N * sizeof(gal_seg_queue));
  const uint64_t o_ring = take(N * B * sizeof(uint32_t));
  h.block_bits = (N * B) <= 1 ? 0 : 64u - (uint32_t)__builtin_clzll(N * B - 1);
  const uint64_t o_blocks = take((1ULL << h.block_bits) * sizeof(gal_block));
  const uint64_t o_slots = take((uint64_t)total * sizeof(uint32_t));
  const uint64_t o_locks = take((1 + (uint64_t)T * P) * sizeof(uint32_t));
  const uint64_t o_stats = take(kStatWords * sizeof(unsigned long long));
  const uint64_t o_trees = take(words * sizeof(uint64_t));
  a->meta_bytes = off;
  hipError_t e = hipMalloc(&a->meta, off);
  if (e != hipSuccess) {
    delete a;
    return hip_fail(e, "hipMalloc(metadata)");
  }
  a->pool_bytes = N * S;
  a->pool_kind = (cfg->flags & GAL_FLAG_POOL_HOST) ? 1 : (cfg->flags & GAL_FLAG_POOL_MANAGED) ? 2 : 0;
  e = alloc_pool(a);
  if (e != hipSuccess) {
    (void)hipFree(a->meta);
    delete a;
    return hip_fail(e, "pool allocation");
  }
  char* base = static_cast<char*>(a->meta);
  h.memory = static_cast<char*>(a->pool);
  h.seg_state = reinterpret_cast<uint64_t*>(base + o_seg);
  h.seg_queue = reinterpret_cast<gal_seg_queue*>(base + o_queue);
  h.ring = reinterpret_cast<uint32_t*>(base + o_ring);
  h.blocks = reinterpret_cast<gal_block*>(base + o_blocks);
  h.slots = reinterpret_cast<uint32_t*>(base + o_slots);
  h.locks = reinterpret_cast<uint32_t*>(base + o_locks);
  h.stats = reinterpret_cast<unsigned long long*>(base + o_stats);
  uint64_t* tree_base = reinterpret_cast<uint64_t*>(base + o_trees);
  a->tree_base = tree_base;
  h.seg_tree.words = tree_base + reinterpret_cast<uint64_t>(h.seg_tree.words);
  for (uint32_t t = 0; t < T * P; ++t)
    h.sub_trees[t].words = tree_base + reinterpret_cast<uint64_t>(h.sub_trees[t].words);
  a->dev = reinterpret_cast<gal_state*>(base + o_state);
  h.table = reinterpret_cast<memory_table*>(a->dev);

  // the pool starts zeroed (memory_table.cuh:518; a host pool is zeroed by
  // the host, memory_table.cuh:537)
  int rc = GAL_OK;
  if (a->pool_kind == 1)
    std::memset(a->pool, 0, a->pool_bytes);
  else if ((e = hipMemset(a->pool, 0, a->pool_bytes)) != hipSuccess)
    rc = hip_fail(e, "hipMemset(pool)");
  if (rc == GAL_OK && (e = hipMemset(a->meta, 0, a->meta_bytes)) != hipSuccess)
    rc = hip_fail(e, "hipMemset(metadata)");
  if (rc == GAL_OK && (e = hipMemset(h.slots, 0xFF, (size_t)total * 4)) != hipSuccess)
    rc = hip_fail(e, "hipMemset(slots)");
  if (rc == GAL_OK && (e = hipMemcpy(a->dev, &h, sizeof(h), hipMemcpyHostToDevice)) != hipSuccess)
    rc = hip_fail(e, "hipMemcpy(state)");
  if (rc == GAL_OK) {
    k_init_segments<<<(unsigned)((N + 255) / 256), 256>>>(a->dev);
    k_init_blocks<<<(unsigned)((N * B + 255) / 256), 256>>>(a->dev);
    const uint64_t top_words = h.seg_tree.layer_words[h.seg_tree.num_layers - 1];
    k_fill_tree<<<(unsigned)((top_words + 255) / 256), 256>>>(h.seg_tree);
    // with GAL_FLAG_LAZY_BOOT the slots start empty and each is filled by
    // the first request that finds it so (the refill path)
    if (!(cfg->flags & GAL_FLAG_LAZY_BOOT)) k_boot<<<1, 64>>>(a->dev);
    if ((e = hipDeviceSynchronize()) != hipSuccess) rc = hip_fail(e, "boot");
  }
  if (rc == GAL_OK) {
    unsigned long long boot_failed = 0;
    const hipError_t e = hipMemcpy(&boot_failed, h.stats + kStatBootFailed,
                                   sizeof(boot_failed), hipMemcpyDeviceToHost);
    if (e != hipSuccess)
      rc = hip_fail(e, "read boot status");
    else if (boot_failed)
      rc = fail(GAL_ERR_BOOT, "pool too small for the pinned boot blocks");
  }
  if (rc != GAL_OK) {
    (void)free_pool(a);
    (void)hipFree(a->meta);
    delete a;
    return rc;
  }
  {
    std::lock_guard<std::mutex> g(g_registry_mu);
    g_registry[a->dev] = a;
  }
  if (cfg->print_info)
    printf("Booted Gallatin with %u trees in range %llu-%llu and %f GB of %smemory %llu segments\n",
           T, (unsigned long long)m, (unsigned long long)M,
           1.0 * (double)cfg->max_bytes / 1024 / 1024 / 1024,
           a->pool_kind == 1 ? "pinned host " : a->pool_kind == 2 ? "managed " : "",
           (unsigned long long)N);
  *out = a;
  return GAL_OK;
}

int gal_destroy(gal_allocator* a) {
  if (!a) return fail(GAL_ERR_INVALID, "null allocator");
  {
    std::lock_guard<std::mutex> g(g_registry_mu);
    g_registry.erase(a->dev);
  }
  hipError_t e1 = hipDeviceSynchronize();
  hipError_t e2 = free_pool(a);
  hipError_t e3 = hipFree(a->meta);
  delete a;
  if (e1 != hipSuccess) return hip_fail(e1, "hipDeviceSynchronize");
  if (e2 != hipSuccess) return hip_fail(e2, "hipFree(pool)");
  if (e3 != hipSuccess) return hip_fail(e3, "hipFree(metadata)");
  return GAL_OK;
}

void* gal_device_state(gal_allocator* a) { return a ? a->dev : nullptr; }

gal_allocator* gal_from_device_state(const void* dev_state) {
  std::lock_guard<std::mutex> g(g_registry_mu);
  auto it = g_registry.find(dev_state);
  return it == g_registry.end() ? nullptr : it->second;
}

int gal_drain(gal_allocator* a, void* stream) {
  if (!a) return fail(GAL_ERR_INVALID, "null allocator");
  k_drain<<<1, 64, 0, (hipStream_t)stream>>>(a->dev);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? GAL_OK : hip_fail(e, "drain launch");
}

int gal_reboot(gal_allocator* a, void* stream) {
  if (!a) return fail(GAL_ERR_INVALID, "null allocator");
  if (a->host.flags & GAL_FLAG_LAZY_BOOT) return GAL_OK;  // slots fill on demand
  k_boot<<<1, 64, 0, (hipStream_t)stream>>>(a->dev);
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "reboot");
  unsigned long long boot_failed = 0;
  (void)hipMemcpy(&boot_failed, a->host.stats + kStatBootFailed, sizeof(boot_failed), hipMemcpyDeviceToHost);
  if (boot_failed) return fail(GAL_ERR_BOOT, "reboot ran out of segments");
  return GAL_OK;
}

int gal_snapshot(gal_allocator* a, uint64_t* seg_state, uint32_t* slots, uint32_t* blocks) {
  if (!a) return fail(GAL_ERR_INVALID, "null allocator");
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "hipDeviceSynchronize");
  const gal_state& h = a->host;
  if (seg_state && (e = hipMemcpy(seg_state, h.seg_state, h.num_segments * 8, hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_fail(e, "snapshot seg_state");
  if (slots && (e = hipMemcpy(slots, h.slots, (size_t)h.total_slots * 4, hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_fail(e, "snapshot slots");
  if (blocks) {
    // gather block counters back into block-id order
    std::vector<gal_block> raw(1ULL << h.block_bits);
    if ((e = hipMemcpy(raw.data(), h.blocks, raw.size() * sizeof(gal_block), hipMemcpyDeviceToHost)) != hipSuccess)
      return hip_fail(e, "snapshot blocks");
    const uint64_t nblk = h.num_segments << h.bps_shift;
    for (uint64_t i = 0; i < nblk; ++i) {
      const gal_block& b = raw[block_position(i, h.block_bits)];
      blocks[2 * i] = b.malloc_counter;
      blocks[2 * i + 1] = b.free_counter;
    }
  }
  return GAL_OK;
}

int64_t gal_tree_words(gal_allocator* a, int tree, int partition, uint64_t* out,
                       uint64_t cap, uint32_t* num_layers, uint64_t* layer_words) {
  if (!a) return fail(GAL_ERR_INVALID, "null allocator");
  const gal_state& h = a->host;
  if (tree < -1 || tree >= (int)h.num_trees) return fail(GAL_ERR_INVALID, "bad tree index");
  if (partition < 0 || partition >= (int)h.num_parts)
    return fail(GAL_ERR_INVALID, "bad partition index");
  const gal_tree& t = tree < 0 ? h.seg_tree : h.sub_trees[tree * h.num_parts + partition];
  const uint64_t n = tree_total_words(t);
  if (num_layers) *num_layers = t.num_layers;
  if (layer_words)
    for (uint32_t l = 0; l < t.num_layers; ++l) layer_words[l] = t.layer_words[l];
  if (!out) return (int64_t)n;
  if (cap < n) return fail(GAL_ERR_INVALID, "buffer too small");
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(out, t.words, n * 8, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "tree copy");
  return (int64_t)n;
}

int gal_get_info(gal_allocator* a, gal_info* out) {
  if (!a || !out) return fail(GAL_ERR_INVALID, "null argument");
  const gal_state& h = a->host;
  std::memset(out, 0, sizeof(*out));
  out->num_segments = h.num_segments;
  out->bytes_per_segment = h.bytes_per_segment;
  out->smallest = h.smallest;
  out->biggest = h.biggest;
  out->blocks_per_segment = h.blocks_per_segment;
  out->num_trees = h.num_trees;
  out->total_slots = h.total_slots;
  out->partitions = h.num_parts;
  out->pool_base = (uint64_t)h.memory;
  std::vector<uint64_t> words(a->tree_words);
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(words.data(), a->tree_base, a->tree_words * 8, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "info copy");
  auto fill = [&](const gal_tree& t) {
    const uint32_t L = t.num_layers - 1;
    const uint64_t* w = words.data() + (t.words - a->tree_base) + t.layer_off[L];
    uint64_t c = 0;
    for (uint64_t i = 0; i < t.layer_words[L]; ++i) c += (uint64_t)__builtin_popcountll(w[i]);
    return c;
  };
  out->segments_free = fill(h.seg_tree);
  for (uint32_t t = 0; t < h.num_trees; ++t)
    for (uint32_t p = 0; p < h.num_parts; ++p)
      out->tree_owned[t] += fill(h.sub_trees[t * h.num_parts + p]);
  std::vector<uint64_t> st(h.num_segments);
  if ((e = hipMemcpy(st.data(), h.seg_state, h.num_segments * 8, hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_fail(e, "info copy");
  for (uint64_t i = 0; i < h.num_segments; ++i) {
    const uint32_t t = (uint32_t)(st[i] & 0xFFFF);
    const int32_t active = (int32_t)(st[i] >> 32);
    if (t < h.num_trees) {
      out->segments_formatted++;
      if (active != -1) {  // a closed segment is on its way back (print_guided_fill skips it)
        const uint64_t nb = h.blocks_per_segment >> t;
        const uint64_t avail = active < 0 ? 0 : (uint64_t)active + 1;
        out->class_segments[t]++;
        out->class_blocks_free[t] += avail;
        out->class_blocks_used[t] += avail < nb ? nb - avail : 0;
      }
    } else if (t != kFreeTree) {
      out->segments_multi++;
    }
  }
  unsigned long long stats[kStatCount];
  if ((e = hipMemcpy(stats, h.stats, sizeof(stats), hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_fail(e, "info copy");
  out->errors = stats[kStatErrors];
  out->overhead_bytes = a->meta_bytes;
  return GAL_OK;
}

int gal_get_stats(gal_allocator* a, uint64_t* out, uint32_t n) {
  if (!a || (!out && n)) return fail(GAL_ERR_INVALID, "null argument");
  unsigned long long stats[kStatCount];
  const hipError_t e = hipMemcpy(stats, a->host.stats, sizeof(stats), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "stats copy");
  for (uint32_t i = 0; i < n && i < (uint32_t)kStatCount; ++i) out[i] = stats[i];
  return (int)kStatCount;
}

int gal_print_info(gal_allocator* a) {
  gal_info info;
  int rc = gal_get_info(a, &info);
  if (rc != GAL_OK) return rc;
  // print_info (gallatin.cuh:1855-1897) / print_usage / print_overhead
  printf("\nGallatin usage stats:\n");
  printf("Segment tree sees %llu/%llu segments available\n",
         (unsigned long long)info.segments_free, (unsigned long long)info.num_segments);
  for (uint32_t t = 0; t < info.num_trees; ++t)
    printf("Tree %u: size %llu, owns %llu/%llu\n", t,
           (unsigned long long)(info.smallest << t), (unsigned long long)info.tree_owned[t],
           (unsigned long long)info.num_segments);
  // print_usage / print_guided_fill (gallatin.cuh:1946-1990): per class, the
  // segments formatted for it and their average free and handed-out blocks
  printf("\nTree usage stats:\n");
  for (uint32_t t = 0; t < info.num_trees; ++t) {
    const uint64_t n = info.class_segments[t];
    const unsigned long long nb = (unsigned long long)(info.blocks_per_segment >> t);
    printf("Tree %u: %llu live segments | avg available %f / %llu | avg in use %f / %llu\n", t,
           (unsigned long long)n, n ? 1.0 * info.class_blocks_free[t] / n : 0.0, nb,
           n ? 1.0 * info.class_blocks_used[t] / n : 0.0, nb);
  }
  printf("Formatted segments %llu, multi-segment heads %llu, errors %llu\n",
         (unsigned long long)info.segments_formatted, (unsigned long long)info.segments_multi,
         (unsigned long long)info.errors);
  printf("Allocator is using %llu bytes of overhead\n", (unsigned long long)info.overhead_bytes);
  fflush(stdout);
  return GAL_OK;
}

}  // extern "C"

// Workgroup-batched malloc / free: the requests of a whole workgroup are
// bucketed in LDS and served together.
//
//   malloc: requests are counted per size class in LDS; one wave then claims
//     every class's whole count at once, lane c for class c, with one
//     counter compare-and-swap per class (more only when it races another
//     claim or a claim runs past slice 4095 and the block is replaced), and
//     each request takes its slice by its rank in the class.  This is the reference's coalesced group claim
//     (gallatin.cuh:865-1078, block.cuh:245-264) with the group widened from
//     a warp to a workgroup.
//   free: requests are grouped per block in an LDS hash table and each block
//     gets one free-counter atomic (gallatin.cuh:1774-1827 with a
//     workgroup-wide team).
//   batch_step does one free and one malloc per thread with two barriers,
//     the class claims of wave 0 running beside the block releases of the
//     other waves.
//
// Requests the batch cannot serve (whole blocks, multi-segment runs, runs of
// largest slices in batch_step or across two claims, or a class that needs
// more than kRanges claims) go through
// the per-wave path of gallatin.hpp, so results and failure behaviour are
// those of malloc / free.  In the deterministic mode the class claims take
// slot 0 and smallest-id segments like the per-wave path, so the slices a
// batch hands out per class are exactly those of one coalesced claim of that
// many consecutive slices (gallatin.cuh:865-1078); which request gets which
// one depends on scheduling.  Frees of the deterministic mode take the
// per-wave path.
//
// Every thread of the workgroup must call these functions (workgroup-uniform
// control flow; blockDim.x a multiple of 64 and at most 1024), and the LDS
// state must be set up once with batch_init before the first call.
#ifndef GALLATIN_BATCHED_HPP
#define GALLATIN_BATCHED_HPP

#include "gallatin.hpp"

namespace gallatin {
namespace allocators {
namespace batch {

constexpr int kRanges = 4;        // claims per class per batch
constexpr int kFreeSlots = 2048;  // LDS hash slots for block grouping

struct lds_state {
  // malloc
  uint32_t count[kMaxTrees];
  uint32_t nranges[kMaxTrees];
  uint32_t blk[kMaxTrees][kRanges];
  uint32_t start[kMaxTrees][kRanges];
  uint32_t len[kMaxTrees][kRanges];
  uint32_t first[kMaxTrees][kRanges];
  // free
  uint32_t key[kFreeSlots];  // block id + 1, 0 = empty
  uint32_t cnt[kFreeSlots];
};

// Clear the LDS state (whole workgroup, once per kernel).
__device__ inline void batch_init(lds_state& L) {
  for (uint32_t i = threadIdx.x; i < (uint32_t)kFreeSlots; i += blockDim.x) L.key[i] = 0, L.cnt[i] = 0;
  for (uint32_t i = threadIdx.x; i < (uint32_t)kMaxTrees; i += blockDim.x) L.count[i] = 0;
  __syncthreads();
}

// Claims of all classes with pending requests, lane c for class c (one
// wave, uniform).  A class whose claim ran out of its block retries in lane
// order until served, out of ranges or out of attempts.
__device__ inline void claim_classes(gal_state* s, lds_state& L, uint32_t p) {
  const uint32_t T = s->num_trees;
  const bool det = detail::deterministic(s);
  const uint32_t c = wave::lane();
  const uint32_t need = c < T ? L.count[c] : 0u;
  if (c < T) L.count[c] = 0;  // ready for the next batch
  const bool mine = need != 0;
  uint32_t served = 0, nr = 0, attempts = 0;
  bool open = mine;
  while (__ballot(open)) {
    uint64_t sb = kNullBlock;
    if (open) sb = detail::pinned_block(s, c, p);
    // classes that find no pinned block refill their start slot, in lane
    // (class) order
    for (uint64_t m = __ballot(open && (uint32_t)sb == kNullBlock); m; m &= m - 1) {
      const int l = wave::first(m);
      const uint32_t sl = (uint32_t)(wave::readlane(sb, l) >> 32);
      const uint32_t got = detail::refill_slot(s, (uint32_t)l, p, sl, det);
      if ((int)c == l) sb = ((uint64_t)sl << 32) | got;
    }
    const uint32_t slot = (uint32_t)(sb >> 32), blk = (uint32_t)sb;
    const bool have = open && blk < kRefilling;
    const bool waiting = open && blk == kRefilling;  // another wave refills
    // The claim takes exactly what the block has left, by compare-and-swap,
    // and never adds to a full or stale block.  A workgroup's demand can be
    // many blocks' worth (runs of largest slices over a whole chunk), and
    // adding it to a full block that is about to be replaced, once per
    // retrying workgroup, would carry the 20-bit count into the tag.
    uint32_t cnt = 0, take = 0;
    bool tag_ok = false;
    if (have) {
      uint32_t* mc = &detail::block_at(s, blk)->malloc_counter;
      uint32_t cur = utils::atomic_load(mc);
      for (uint32_t g = 0; g < 64; ++g) {
        cnt = cur & kCountMask;
        tag_ok = (cur >> kTreeShift) == block_tag(c, p);
        if (!tag_ok || cnt >= 4096u) break;
        const uint32_t want = min(need - served, 4096u - cnt);
        const uint32_t seen = utils::atomic_cas(mc, cur, cur + want);
        if (seen == cur) {
          take = want;
          break;
        }
        cur = seen;
      }
    }
    const bool got = take > 0;
    if (got) {
      L.blk[c][nr] = blk;
      L.start[c][nr] = cnt;
      L.len[c][nr] = take;
      L.first[c][nr] = served;
      ++nr;
      served += take;
    }
    // the claim that took slice 4095 replaces the block, one class at a
    // time; so does a class that keeps finding its block full (its 4095
    // claimer may have detached nothing); the slot CAS admits one of them
    const uint64_t rep = __ballot((got && cnt + take == 4096u) ||
                                  (tag_ok && cnt >= 4096u && attempts >= kHelpAfter));
    for (uint64_t m = rep; m; m &= m - 1) {
      const int l = wave::first(m);
      detail::replace_block(s, wave::readlane(c, l), p, wave::readlane(slot, l),
                            wave::readlane(blk, l), det);
    }
    if (__ballot(waiting)) utils::backoff();
    if (open && (have || waiting) &&
        !(served == need || nr == (uint32_t)kRanges || ++attempts >= s->lim_slice))
      continue;
    open = false;
  }
  if (c < T) L.nranges[c] = nr;
}

// ---- malloc side ----

struct malloc_req {
  uint32_t t;     // size class
  uint32_t rank;  // first slice among the workgroup's claims of class t
  uint32_t cnt;   // slices: 1, or a run of slices of the largest class
  bool batched;   // served by the class claims (else the wave path)
};

// Class and weight of a request, without its rank.  With kRuns, a request
// above the largest slice that malloc serves as a run of cnt largest slices
// (gallatin.cuh:1210-1217) joins the largest class's claim with weight cnt;
// the deterministic mode leaves runs to the wave path.
template <bool kRuns = false>
__device__ __forceinline__ malloc_req malloc_classify(gal_state* s, bool want, uint64_t size) {
  malloc_req r{0u, 0u, 1u, false};
  if (want) {
    const int cl = utils::get_first_bit_bigger(size);
    r.t = size < s->smallest ? 0u : (uint32_t)(cl - (int)s->smallest_bits);
    r.batched = r.t < s->num_trees;
    if (kRuns && !r.batched && cl < (int)s->block_tree_bits && !detail::deterministic(s)) {
      r.cnt = 1u << (r.t - (s->num_trees - 1));
      r.t = s->num_trees - 1;
      r.batched = true;
    }
  }
  return r;
}

// Phase 1: class and rank of this thread's request (malloc_classify, then
// its place in the workgroup's count of its class).
template <bool kRuns = false>
__device__ __forceinline__ malloc_req malloc_stage(gal_state* s, bool want, uint64_t size,
                                                   lds_state& L) {
  malloc_req r = malloc_classify<kRuns>(s, want, size);
  if (r.batched) r.rank = atomicAdd(&L.count[r.t], r.cnt);
  return r;
}

// Phase 3: the byte offset of this thread's slice, or kFail.  The ranges
// are rewritten only by the next batch's claims, after a barrier every
// thread reaches after reading them here.  A run takes its slices only when
// they lie in one range, that is in one block, and then credits the cnt - 1
// frees its one free will not make (block_correct_frees, block.cuh:267-281).
// A run across two ranges, or past what the claims served, credits back
// the slices the claims took for it and goes to the wave path.
template <bool kRuns = false>
__device__ __forceinline__ uint64_t malloc_take(gal_state* s, const malloc_req& r,
                                                const lds_state& L, uint32_t p) {
  if (!r.batched) return kFail;
  const uint32_t nr = L.nranges[r.t];
  for (uint32_t k = 0; k < nr; ++k) {
    const uint32_t f = L.first[r.t][k];
    if (r.rank >= f && r.rank < f + L.len[r.t][k]) {
      if (kRuns && r.rank + r.cnt > f + L.len[r.t][k]) break;
      if (kRuns && r.cnt > 1) detail::block_release(s, L.blk[r.t][k], r.cnt - 1, p);
      const uint64_t off = (uint64_t)L.start[r.t][k] + (r.rank - f) +
                           ((uint64_t)L.blk[r.t][k] << 12);
      return detail::offset_to_bytes(s, off, r.t);
    }
  }
  if (kRuns && r.cnt > 1) {
    for (uint32_t k = 0; k < nr; ++k) {
      const uint32_t f = L.first[r.t][k], e = f + L.len[r.t][k];
      const uint32_t lo = r.rank > f ? r.rank : f;
      const uint32_t hi = r.rank + r.cnt < e ? r.rank + r.cnt : e;
      if (lo < hi) detail::block_release(s, L.blk[r.t][k], hi - lo, p);
    }
  }
  return kFail;
}

// ---- free side ----

// Phase 1: orders this wave's payload before the block releases (agent
// scope if one of its slices lives in another XCD's segment) and groups the
// slice frees per block in the LDS table.  Returns false for a pointer the
// table does not take (multi-segment run, invalid pointer, table overflow),
// which then goes through the wave path.
__device__ __forceinline__ bool free_stage(gal_state* s, void* ptr, lds_state& L,
                                           uint32_t my_p) {
  const uint64_t base = (uint64_t)s->memory;
  const uint64_t pa = (uint64_t)ptr;
  bool slice = false, foreign = false;
  uint32_t blk = 0;
  if (ptr != nullptr && !detail::deterministic(s) &&
      (pa - base) < (s->num_segments << s->seg_shift)) {
    const uint64_t seg = (pa - base) >> s->seg_shift;
    const uint64_t st = utils::atomic_load(&s->seg_state[seg]);
    const uint32_t t = (uint32_t)(st & 0xFFFF);
    if (t < s->num_trees) {
      slice = true;
      blk = (uint32_t)(detail::bytes_to_offset(s, pa - base, t) >> 12);
      foreign = seg_home(st) != my_p;
    }
  }
  if (!(s->flags & kFlagNoPayloadFence) && __ballot(ptr != nullptr)) {
    if (s->num_parts == 1 || __ballot(foreign))
      utils::fence_release();
    else
      detail::release_to_xcd();
  }
  if (!slice) return false;
  uint32_t h = (blk * 0x9E3779B1u) >> 21;  // 11 bits
  for (int probe = 0; probe < 64; ++probe) {
    const uint32_t prev = atomicCAS(&L.key[h], 0u, blk + 1);
    if (prev == 0u || prev == blk + 1) {
      atomicAdd(&L.cnt[h], 1u);
      return true;
    }
    h = (h + 1) & (kFreeSlots - 1);
  }
  return false;
}

// Phase 2: one free-counter atomic per block in the table, by threads
// [first, blockDim) of the workgroup; empties the table.
__device__ __forceinline__ void free_scan(gal_state* s, lds_state& L, uint32_t my_p,
                                          uint32_t first) {
  const uint32_t span = blockDim.x - first;
  for (uint32_t i = threadIdx.x - first; i < (uint32_t)kFreeSlots; i += span) {
    const uint32_t key = L.key[i];
    if (!key) continue;
    const uint32_t b = key - 1, k = L.cnt[i];
    L.key[i] = 0;
    L.cnt[i] = 0;
    const uint32_t old = utils::atomic_add(&detail::block_at(s, b)->free_counter, k);
    if (old + k == kSlicesPerBlock) {
      utils::atomic_exch(&detail::block_at(s, b)->free_counter, 0u);
      detail::free_block(s, b, my_p);
    }
  }
}

// free_scan for a workgroup that nothing allocates from while it frees
// (a cleanup): a block its frees empty is not returned on the spot, where
// the one thread returning it would hold every thread of the workgroup at
// the next barrier; its id goes to `later` (`cap` entries, count in
// *nlater, both in LDS) and the workgroup returns the listed blocks
// together after its last barrier (return_later).  Past the capacity a
// block is returned at once.  An emptied block is no longer pinned, so no
// slice of it is handed out before it is returned.
__device__ __forceinline__ void free_scan_later(gal_state* s, lds_state& L, uint32_t my_p,
                                                uint32_t* later, uint32_t* nlater,
                                                uint32_t cap) {
  for (uint32_t i = threadIdx.x; i < (uint32_t)kFreeSlots; i += blockDim.x) {
    const uint32_t key = L.key[i];
    if (!key) continue;
    const uint32_t b = key - 1, k = L.cnt[i];
    L.key[i] = 0;
    L.cnt[i] = 0;
    const uint32_t old = utils::atomic_add(&detail::block_at(s, b)->free_counter, k);
    if (old + k == kSlicesPerBlock) {
      utils::atomic_exch(&detail::block_at(s, b)->free_counter, 0u);
      const uint32_t at = atomicAdd(nlater, 1u);
      if (at < cap)
        later[at] = b;
      else
        detail::free_block(s, b, my_p);
    }
  }
}

// The blocks free_scan_later listed, returned by the workgroup's threads
// together; call after a barrier that follows the last free_scan_later.
__device__ inline void return_later(gal_state* s, uint32_t my_p, const uint32_t* later,
                                    uint32_t nlater, uint32_t cap) {
  const uint32_t n = nlater < cap ? nlater : cap;
  for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) detail::free_block(s, later[j], my_p);
}

// ---- workgroup entry points ----

// Each thread with `want` gets a pointer of `size` bytes (or nullptr).
// Runs of largest slices are claimed with the classes (see malloc_stage).
__device__ inline void* batch_malloc(gal_state* s, bool want, uint64_t size, lds_state& L) {
  const uint32_t p = detail::wave_part(s);
  const malloc_req r = malloc_stage<true>(s, want, size, L);
  __syncthreads();
  if (threadIdx.x < 64) claim_classes(s, L, p);
  __syncthreads();
  const uint64_t bytes = malloc_take<true>(s, r, L, p);
  void* out = bytes == kFail ? nullptr : (void*)(s->memory + bytes);
  const bool rest = want && bytes == kFail;
  if (__ballot(rest)) {
    void* q = nullptr;
    if (rest) q = detail::malloc_impl(s, size);
    if (rest) out = q;
  }
  return out;  // no acquire: see malloc_impl
}

// Each thread frees `ptr` (nullptr: nothing).
__device__ inline void batch_free(gal_state* s, void* ptr, lds_state& L) {
  const uint32_t my_p = detail::wave_part(s);
  const bool grouped = free_stage(s, ptr, L, my_p);
  __syncthreads();
  free_scan(s, L, my_p, 0);
  const bool rest = ptr != nullptr && !grouped;
  if (__ballot(rest)) {
    if (rest) detail::free_impl(s, ptr);
  }
  __syncthreads();
}

// One step of a workgroup: every thread frees `old_ptr` (nullptr: nothing)
// and, with `want`, gets a pointer of `size` bytes (or nullptr), with two
// barriers in all.  Wave 0 claims the size classes while the other waves
// release the freed blocks.  The new slices never come from the frees of
// the same step (a block returns to service only once all its slices are
// free), so doing both at once changes no result of either.
__device__ inline void* batch_step(gal_state* s, void* old_ptr, bool want, uint64_t size,
                                   lds_state& L) {
  const uint32_t p = detail::wave_part(s);
  const bool grouped = free_stage(s, old_ptr, L, p);
  const malloc_req r = malloc_stage(s, want, size, L);
  __syncthreads();
  if (threadIdx.x < 64) claim_classes(s, L, p);
  free_scan(s, L, p, blockDim.x > 64 ? 64u : 0u);
  __syncthreads();
  const bool rest_free = old_ptr != nullptr && !grouped;
  if (__ballot(rest_free)) {
    if (rest_free) detail::free_impl(s, old_ptr);
  }
  const uint64_t bytes = malloc_take(s, r, L, p);
  void* out = bytes == kFail ? nullptr : (void*)(s->memory + bytes);
  const bool rest = want && bytes == kFail;
  if (__ballot(rest)) {
    void* q = nullptr;
    if (rest) q = detail::malloc_impl(s, size);
    if (rest) out = q;
  }
  return out;  // no acquire: see malloc_impl
}

}  // namespace batch
}  // namespace allocators
}  // namespace gallatin

#endif

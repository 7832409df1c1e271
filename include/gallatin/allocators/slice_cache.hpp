// Per-workgroup slice caches in LDS.
//
// Each workgroup keeps, per size class, one range of slices it has already
// claimed from the class's pinned block, packed in one 64-bit LDS word:
//
//   (block id << 32) | (next slice << 16) | end slice
//
// A wave groups its requests by class, and each group's leader adds the
// group's count to its class's "next" field with one LDS atomic (all
// classes in the same instructions): no barrier, no global atomic.  The one
// group whose add crosses "end" claims a new range from the pinned block for
// its unserved lanes plus a stock for the workgroup (one counter atomic, the
// reference's coalesced claim of gallatin.cuh:865-1078 / block.cuh:245-264
// with the group widened to a stock), hands out its lanes' slices and
// publishes the rest.  Requests that meet an exhausted range before the
// crosser has published go through the per-wave path of gallatin.hpp, as do
// requests the cache does not take (whole blocks, multi-segment runs, runs
// of 4 KiB slices).  A range's unused slices stay
// claimed until cache_flush gives them back to their block at the end of
// the kernel, so every slice is accounted for: after the flush the block
// counters read as if each cached slice had been allocated and freed.
//
// "next" only grows while a range is live, by at most one add per wave
// and class that saw next <= end before adding (at most 16 waves x 64 lanes
// past end), so the 16-bit field never carries into the block id.
//
// Frees are not cached: they take the per-wave path (block-grouped counter
// atomics), which already issues one atomic per block per wave.
#ifndef GALLATIN_SLICE_CACHE_HPP
#define GALLATIN_SLICE_CACHE_HPP

#include "gallatin.hpp"

namespace gallatin {
namespace allocators {
namespace cache {

struct slice_cache {
  unsigned long long w[kMaxTrees];
};

constexpr unsigned long long kEmpty = (unsigned long long)kNullBlock << 32;  // next = end = 0

// Stock a crossing wave claims beyond its own lanes' needs, in slices of
// class t: 256 slices of 16 B down to 32 of 4 KiB (at most 128 KiB per
// class per workgroup).
__device__ __forceinline__ uint32_t stock_of(uint32_t t) {
  const uint32_t k = 256u >> (t > 3 ? 3 : t);
  return k;
}

// Clear the cache (whole workgroup, once per kernel; ends with a barrier).
__device__ inline void cache_init(slice_cache& L) {
  for (uint32_t i = threadIdx.x; i < (uint32_t)kMaxTrees; i += blockDim.x) L.w[i] = kEmpty;
  __syncthreads();
}

// One range of up to `want` slices of class t (want <= 4096) from the
// pinned block of partition p (wave op, wave-uniform arguments).  Returns
// (blk << 32) | (first slice << 16) | count, count 0 when none could be
// had.  The claim, the stale-block check and the block replacement follow
// malloc_slices.
__device__ inline uint64_t claim_range(gal_state* s, uint32_t t, uint32_t p, uint32_t want) {
  const bool det = detail::deterministic(s);
  for (uint32_t attempts = 0;;) {
    uint64_t sb = kNullBlock;
    if (wave::is_first()) sb = detail::pinned_block(s, t, p);
    sb = wave::readfirst(sb);
    const uint32_t slot = (uint32_t)(sb >> 32);
    uint32_t blk = (uint32_t)sb;
    if (blk == kNullBlock) blk = detail::refill_slot(s, t, p, slot, det);
    if (blk == kNullBlock) return 0;  // no block to be had
    if (blk != kRefilling) {
      uint32_t old = 0;
      if (wave::is_first()) old = utils::atomic_add(&detail::block_at(s, blk)->malloc_counter, want);
      old = wave::readfirst(old);
      const uint32_t cnt = old & kCountMask;
      const bool tag_ok = (old >> kTreeShift) == block_tag(t, p);
      const uint32_t take = cnt < kSlicesPerBlock ? min(want, kSlicesPerBlock - cnt) : 0u;
      if (!tag_ok && take && wave::is_first()) detail::block_release(s, blk, take, p);
      // the claim that took slice 4095 replaces the block; so does a wave
      // that keeps finding its class's block full
      if (tag_ok && ((take && cnt + take == kSlicesPerBlock) ||
                     (cnt >= kSlicesPerBlock && attempts >= kHelpAfter)))
        detail::replace_block(s, t, p, slot, blk, det);
      if (tag_ok && take) return ((uint64_t)blk << 32) | ((uint64_t)cnt << 16) | take;
    } else {
      utils::backoff();  // another wave refills the slot
    }
    if (++attempts >= s->lim_slice) return 0;
  }
}

// malloc through the cache (every lane of the wave calls it; `want` says
// whether this lane has a request).  nullptr for lanes without one and for
// failed requests.  The lanes are grouped by class and every group's leader
// reads and adds to its class's word in the same instructions, so a wave
// with requests of eight classes pays two LDS round trips, not sixteen;
// only groups that cross the end of their range claim, one after another.
__device__ inline void* cache_malloc(gal_state* s, slice_cache& L, bool want, uint64_t size) {
  const uint32_t T = s->num_trees;
  const uint32_t lane = wave::lane();
  const int cl = utils::get_first_bit_bigger(size < 1 ? 1 : size);
  const uint32_t t = size < s->smallest ? 0u : (uint32_t)(cl - (int)s->smallest_bits);
  const bool cached = want && t < T;
  const uint32_t p = detail::wave_part(s);
  uint64_t off = kFail;  // slice offset of this lane's allocation
  const uint64_t pend = __ballot(cached);
  if (pend) {
    const uint64_t grp = wave::match(pend, cached, t);  // lanes of this lane's class
    const int leader = cached ? wave::first(grp) : (int)lane;
    const bool is_leader = cached && leader == (int)lane;
    const uint32_t g = wave::count(grp);
    const uint32_t rank = wave::count(grp & wave::lanes_below());
    unsigned long long old = 0;
    uint32_t added = 0;
    if (is_leader) {
      const unsigned long long cur = __atomic_load_n(&L.w[t], __ATOMIC_RELAXED);
      if (((cur >> 16) & 0xFFFFu) <= (cur & 0xFFFFu)) {
        old = atomicAdd(&L.w[t], (unsigned long long)g << 16);
        added = 1;
      }
    }
    // every lane takes its leader's result (an exhausted range sends the
    // group to the per-wave path)
    added = wave::fetch(added, leader);
    old = wave::fetch((uint64_t)old, leader);
    const bool live = cached && added;
    const uint32_t blk = (uint32_t)(old >> 32);
    const uint32_t next = (uint32_t)(old >> 16) & 0xFFFFu;
    // slices left in the range when the add landed; negative once another
    // wave has crossed its end.  A signed difference on purpose: with
    // `next < end ? end - next : 0` the compiler hoisted the unsigned
    // difference with a no-wrap flag and dropped the next <= end half of
    // the crossing test below, so two waves crossed one range.
    const int32_t room = (int32_t)(old & 0xFFFFu) - (int32_t)next;
    const uint32_t taken = live && room > 0 ? min(g, (uint32_t)room) : 0u;
    if (live && rank < taken) off = ((uint64_t)blk << 12) + next + rank;
    // groups whose add crossed the end claim for their other lanes plus a
    // stock, one group at a time (lane order of the leaders)
    const bool crossed = live && room >= 0 && room < (int32_t)g;
    for (uint64_t m = __ballot(is_leader && crossed); m; m &= m - 1) {
      const int l = wave::first(m);
      const uint32_t c = wave::readlane(t, l);
      const uint64_t gl = wave::readlane(grp, l);
      const uint32_t tl = wave::readlane(taken, l);
      const uint32_t need = wave::count(gl) - tl;
      const uint64_t r = claim_range(s, c, p, need + stock_of(c));
      const uint32_t rb = (uint32_t)(r >> 32);
      const uint32_t rs = (uint32_t)(r >> 16) & 0xFFFFu;
      const uint32_t rl = (uint32_t)r & 0xFFFFu;
      const uint32_t use = min(need, rl);
      if (((gl >> lane) & 1) && rank >= tl && rank - tl < use)
        off = ((uint64_t)rb << 12) + rs + (rank - tl);
      const unsigned long long nw =
          rl > use ? ((unsigned long long)rb << 32) | ((unsigned long long)(rs + use) << 16) |
                         (rs + rl)
                   : kEmpty;
      if ((int)lane == l) atomicExch(&L.w[c], nw);
    }
  }
  const bool got = off != kFail;  // no acquire: see malloc_impl
  void* res = got ? (void*)(s->memory + detail::offset_to_bytes(s, off, t)) : nullptr;
  if (want && !got) res = detail::malloc_impl(s, size);
  return res;
}

// Give every cached range's unused slices back to their blocks (whole
// workgroup, after its last cache_malloc; starts with a barrier).
__device__ inline void cache_flush(gal_state* s, slice_cache& L) {
  __syncthreads();
  if (threadIdx.x < (uint32_t)s->num_trees) {
    const unsigned long long w = L.w[threadIdx.x];
    const uint32_t blk = (uint32_t)(w >> 32);
    const uint32_t next = (uint32_t)(w >> 16) & 0xFFFFu;
    const uint32_t end = (uint32_t)w & 0xFFFFu;
    if (blk != kNullBlock && next < end)
      detail::block_release(s, blk, end - next, detail::wave_part(s));
    L.w[threadIdx.x] = kEmpty;
  }
}

}  // namespace cache
}  // namespace allocators
}  // namespace gallatin

#endif

// Gallatin device allocator for MI355X (gfx950): device-side malloc / free
// over segments, blocks of 4096 slices and multi-segment runs.
//
// The allocation policy is the reference's (saltsystemslab/gallatin,
// include/gallatin/allocators/gallatin.cuh): size classes from `smallest` to
// `biggest` in powers of two, one bitmap tree per class plus the segment
// tree, 4096-slice blocks claimed by one counter atomic, pinned blocks per
// class, recycled blocks in a per-segment ring, segments returned when all
// their blocks are back.  The execution is wave64- and XCD-native:
//
//  * every entry point is called by the active lanes of a wave together; the
//    lanes are partitioned by size class (malloc) or block (free) with
//    ballots, every group's counter atomic is issued in one instruction by
//    the group leaders, and the rare paths (block turnover, segment setup and
//    release, whole-block and multi-segment requests) run one request at a
//    time in lane order with the whole wave available for tree searches;
//  * tree searches scan 64 words per pass across the wave (veb.hpp);
//  * the pinned slots and class trees are split into partitions, one per
//    XCD.  A formatted segment has a home partition; only waves of that XCD
//    pin its blocks, so its memory is written through one L2.  A free on the
//    home XCD therefore only has to wait for its own stores to reach that L2;
//    a free from another XCD (or of a multi-segment run) writes the L2 back
//    (agent-scope release).  A segment changes home only when a wave of its
//    home XCD releases it, after writing its L2 back.
//  * With one partition (deterministic mode, or partitions = 1) every free
//    is an agent-scope release.  A malloc needs no acquire in either case
//    (malloc_impl says why).
//
// Device API (same names as the reference): Gallatin<S,m,M>::malloc / free /
// get_allocated_size / owns_allocation; host API: generate_on_device /
// free_on_device / print_info, implemented over the C ABI in gallatin_capi.h.
#ifndef GALLATIN_HPP
#define GALLATIN_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../gallatin_capi.h"
#include "alloc_utils.hpp"
#include "layout.hpp"
#include "murmurhash.hpp"
#include "veb.hpp"

namespace gallatin {
namespace allocators {
namespace detail {

using utils::atomic_add;
using utils::atomic_and;
using utils::atomic_cas;
using utils::atomic_exch;
using utils::atomic_load;
using utils::atomic_or;
using utils::atomic_store;
using utils::atomic_sub;

enum : uint32_t { kPathSlice = 1, kPathBlock = 2, kPathSegment = 3 };

__device__ __forceinline__ bool deterministic(const gal_state* s) {
  return (s->flags & kFlagDeterministic) != 0;
}

// Partition of the calling wave: its XCD (__smid() = xcc << 6 | se << 4 | cu).
__device__ __forceinline__ uint32_t wave_part(const gal_state* s) {
  return s->num_parts == 1 ? 0u : (__smid() >> 6) % s->num_parts;
}

__device__ __forceinline__ const gal_tree& class_tree(const gal_state* s, uint32_t t,
                                                      uint32_t p) {
  return s->sub_trees[t * s->num_parts + p];
}

// blocks of class t per segment (memory_table.cuh:1057-1061)
__device__ __forceinline__ uint64_t nb_of(const gal_state* s, uint32_t t) {
  return s->blocks_per_segment >> t;
}

// memory_table.cuh:1072-1086 offset_to_allocation, as a byte offset.
__device__ __forceinline__ uint64_t offset_to_bytes(const gal_state* s,
                                                    uint64_t off, uint32_t t) {
  const uint32_t per_seg_shift = s->seg_shift - s->smallest_bits;
  const uint64_t seg = off >> per_seg_shift;
  const uint64_t rel = off & ((1ULL << per_seg_shift) - 1);
  return (seg << s->seg_shift) + (rel << (s->smallest_bits + t));
}

// memory_table.cuh:1090-1127 allocation_to_offset, from a byte offset.
__device__ __forceinline__ uint64_t bytes_to_offset(const gal_state* s,
                                                    uint64_t bytes, uint32_t t) {
  const uint64_t seg = bytes >> s->seg_shift;
  const uint64_t in_seg = bytes & ((1ULL << s->seg_shift) - 1);
  return (in_seg >> (s->smallest_bits + t)) +
         (seg << (s->seg_shift - s->smallest_bits));
}

__device__ __forceinline__ gal_block* block_at(const gal_state* s, uint32_t blk) {
  return s->blocks + block_position(blk, s->block_bits);
}

__device__ __forceinline__ void add_stat(gal_state* s, int idx, uint64_t v) {
  atomic_add(&s->stats[idx], (unsigned long long)v);
}
__device__ __forceinline__ void add_error(gal_state* s, int kind, uint64_t v = 1) {
  add_stat(s, kind, v);
  add_stat(s, kStatErrors, v);
}

// Stores of this wave have reached the L2 of its XCD.
__device__ __forceinline__ void release_to_xcd() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0);
}

// Chunk id field of seg_state: tree 0xFFFF, home 0 (free segment).
__device__ inline void clear_chunk_id(gal_state* s, uint64_t seg) {
  uint64_t cur = atomic_load(&s->seg_state[seg]);
  for (int g = 0; g < 4096; ++g) {
    const uint64_t want = (cur & ~kSegIdMask) | kFreeTree;
    const uint64_t seen = atomic_cas(&s->seg_state[seg], cur, want);
    if (seen == cur) return;
    cur = seen;
  }
  add_error(s, kStatErrChunkId);
}

// ----------------------------------------------------------- per lane ops

// Ring tickets [f, f + k) of segment seg have just been completed: their k
// blocks become available to get_block, and the segment goes back to its
// class tree or, when every block is back, to the segment tree
// (gallatin.cuh:1646-1722, memory_table.cuh:1226-1241).
__device__ inline void complete_run(gal_state* s, uint64_t seg, uint32_t t, uint32_t h,
                                    uint32_t my_p, uint64_t nb, uint32_t k) {
  // REREGISTER_CUTOFF 0.1 (gallatin.cuh:1646) on the segment's fill: returns
  // that bring its free blocks to at least ceil(nb / 10), without completing
  // it, put it back in its class tree when it is not there.  The reference
  // compares the ring position mod nb instead ("TODO - make this actually
  // calculate fill", memory_table.cuh:1181), which hides a segment emptied
  // at ring position k until its ring wraps back to 10 %, about half of its
  // blocks later.  The insert comes before the count increment, so the
  // segment cannot be released in between.
  {
    const gal_tree& ct = class_tree(s, t, h);
    if (!veb::query(ct, seg)) {
      const int64_t f_after =
          (int64_t)(int32_t)(atomic_load(&s->seg_state[seg]) >> 32) + 1 + k;
      if (10 * f_after >= (int64_t)nb && f_after < (int64_t)nb) veb::insert(ct, seg);
    }
  }
  const uint64_t old = atomic_add(&s->seg_state[seg], (uint64_t)k * kActiveOne);
  const int32_t rid = (int32_t)(old >> 32);
  if ((int64_t)rid + k == (int64_t)nb - 1) {
    if (my_p != h) {
      // every block is back but this wave is not on the home XCD, whose L2
      // may still hold lines of the segment: keep it formatted for its home
      veb::insert(class_tree(s, t, h), seg);
      return;
    }
    const uint64_t full = ((uint64_t)(uint32_t)(nb - 1) << 32) | (old & kSegIdMask);
    const uint64_t closed = (0xFFFFFFFFULL << 32) | (old & kSegIdMask);
    if (atomic_cas(&s->seg_state[seg], full, closed) == full) {
      // all blocks back: release the segment (gallatin.cuh:1680-1722)
      veb::remove(class_tree(s, t, h), seg);
      utils::fence_release();
      clear_chunk_id(s, seg);
      utils::fence_release();
      veb::insert(s->seg_tree, seg);
      add_stat(s, kStatSegmentsReleased, 1);
    }
  }
}

// Segment ring return, gallatin.cuh:1637-1726 return_block with
// memory_table.cuh:1158-1241 reserve_segment_slot / finish_freeing_block.
// `h` is the segment's home partition, `my_p` the caller's.
//
// Tickets still complete in order (the free count never covers a ring
// entry that is not written yet), but returners do not wait for each other.
// The reference's returner spins until q_final reaches its own ticket and
// then advances it by one (finalize_free_queue, memory_table.cuh:906-910),
// so returns into one segment form a chain of dependent compare-and-swaps.
// Here the completion state is one word, q_done:
//   bits 63..56  generation, bumped by every set-up of the segment
//   bits 55..32  frontier: the first incomplete ticket, mod 2^24
//   bits 31..0   written-mask over tickets frontier .. frontier + 31,
//                ticket x at bit x mod 32
// A returner writes its entry, then sets its ticket's bit with one atomic
// or.  If the frontier's bit is then set, it moves the frontier over the
// whole run of set bits, clearing them, with one compare-and-swap, and makes
// the run available.  Otherwise it leaves: the frontier only ever stops on
// a clear bit, so the frontier's bit is set by that ticket's own returner,
// after the frontier got there, and that returner sees it set and moves the
// frontier over every set bit after it.  A returner also leaves once the
// frontier has passed its own ticket, so nothing touches the word of a
// segment after its last block came back (the generation makes a
// compare-and-swap from before a new set-up fail).  Only a ticket 32 or
// more ahead of the frontier waits, for the frontier to move.  Marking with
// an or, not a compare-and-swap, lets the lanes of a wave that return into
// one segment mark in one round.
constexpr uint32_t kDoneTickets = 0xFFFFFFu;
constexpr int kDoneGenShift = 56;

__device__ __noinline__ void return_block(gal_state* s, uint32_t blk, uint64_t seg,
                                          uint32_t t, uint32_t h, uint32_t my_p) {
  const uint64_t nb = nb_of(s, t);
  gal_seg_queue* q = &s->seg_queue[seg];
  const uint32_t pos = atomic_add(&q->q_free, 1u);
  atomic_exch(&s->ring[(seg << s->bps_shift) + (pos & (uint32_t)(nb - 1))], blk);
  release_to_xcd();  // the entry is written before its ticket is marked
  // One loop waits for reach, marks this ticket and completes runs, so that
  // lanes of one wave returning into the same segment help each other: a
  // lane waiting for the frontier to come within reach iterates beside the
  // lane that moves it (two loops in a row would hold the mover at the end
  // of the first until the waiter gave up).
  uint64_t w = atomic_load(&q->q_done);
  bool marked = false;
  for (uint32_t spin = 0;; ++spin) {
    const uint32_t f = (uint32_t)(w >> 32) & kDoneTickets;
    // distance of this ticket from the frontier; from 32 on after marking
    // it means the frontier has passed it
    const uint32_t ahead = (pos - f) & kDoneTickets;
    if (!marked) {
      if (ahead >= 32) {
        if (spin >= kSpinCap) {  // the frontier never came within reach
          add_error(s, kStatErrRingOrder);
          return;
        }
        utils::backoff();
        w = atomic_load(&q->q_done);
        continue;
      }
      // Within reach, the ticket's bit (pos mod 32) is free: the frontier
      // only moves forward, and moving it over ticket pos - 32 cleared it.
      const uint64_t bit = (uint64_t)1 << (pos & 31);
      w = atomic_or(&q->q_done, bit) | bit;
      marked = true;
      continue;
    }
    if (ahead >= 32) return;  // another returner completed this ticket
    // the window's bits rotated so that the frontier's comes first
    const uint32_t r = f & 31, bits = (uint32_t)w;
    const uint32_t rot = r ? (bits >> r) | (bits << (32 - r)) : bits;
    if (!(rot & 1u)) return;  // the frontier ticket's returner takes over
    const uint32_t k = rot == 0xFFFFFFFFu ? 32u : (uint32_t)__builtin_ctz(~rot);
    const uint32_t run = k == 32 ? 0xFFFFFFFFu : (1u << k) - 1u;
    const uint32_t cleared = r ? (run << r) | (run >> (32 - r)) : run;
    const uint64_t next = (w & ~(((uint64_t)1 << kDoneGenShift) - 1)) |
                          ((uint64_t)((f + k) & kDoneTickets) << 32) |
                          (uint64_t)(bits & ~cleared);
    const uint64_t seen = atomic_cas(&q->q_done, w, next);
    if (seen != w) {
      w = seen;  // another returner moved the frontier or set a bit; look again
      continue;
    }
    if (k > 1) {
      add_stat(s, kStatRingRuns, 1);
      add_stat(s, kStatRingRunTickets, k);
    }
    // the run stopped at a clear bit (or cleared the whole window), so the
    // new frontier's bit is clear: its own returner will find it
    complete_run(s, seg, t, h, my_p, nb, k);
    return;
  }
}

// gallatin.cuh:1732-1766 free_block.
__device__ __noinline__ void free_block(gal_state* s, uint32_t blk, uint32_t my_p) {
  const uint64_t seg = (uint64_t)blk >> s->bps_shift;
  const uint64_t st = atomic_load(&s->seg_state[seg]);
  const uint32_t t = (uint32_t)(st & 0xFFFF);
  if (t >= s->num_trees) {
    add_error(s, kStatErrBlockFree);
    return;
  }
  return_block(s, blk, seg, t, seg_home(st), my_p);
}

// block.cuh:87-98 block_free_multiple + 200-212 reset_free as used by
// gallatin.cuh:1774-1827 free_offset.
__device__ __forceinline__ void block_release(gal_state* s, uint32_t blk, uint32_t k,
                                              uint32_t my_p) {
  const uint32_t old = atomic_add(&block_at(s, blk)->free_counter, k);
  if (old + k == kSlicesPerBlock) {
    atomic_exch(&block_at(s, blk)->free_counter, 0u);
    free_block(s, blk, my_p);
  }
}

// Give back a block that was handed out but never pinned (it lost a slot
// race).  It is retired the way a full block is: this lane claims every
// slice not yet claimed and credits them as freed.  A wave that read the
// block's id from a slot long ago may still add to its counter; whatever
// it took is freed by that wave, and the block returns to its segment only
// after that.  So a block waiting in a segment ring never shows claimable
// slices (a returned block with count 0 would hand slices of a free segment
// to such a wave).  Per lane.
__device__ inline void retire_block(gal_state* s, uint32_t blk, uint32_t my_p) {
  const uint32_t old = atomic_add(&block_at(s, blk)->malloc_counter, kSlicesPerBlock);
  const uint32_t c = old & kCountMask;
  if (c < kSlicesPerBlock) block_release(s, blk, kSlicesPerBlock - c, my_p);
}

// memory_table.cuh:915-993 get_block for a segment of class t, home p.  The
// packed segment word gives the active count and the owner in one atomic.
// Returns the block id, with bit 32 set when it was the segment's last.
__device__ inline uint64_t get_block(gal_state* s, uint64_t seg, uint32_t t, uint32_t p,
                                    bool look) {
  // With `look`, look before taking: a claim subtracts from the count and
  // adds back when it found none, so with many waves retrying on a segment
  // that has one or two free blocks left every claimer can see the others'
  // transient subtractions and fail, on and on.  A retrying claimer that
  // sees no free block (or another owner) leaves without touching the
  // count.  A first attempt takes directly, which saves the load.
  if (look) {
    const uint64_t cur = atomic_load(&s->seg_state[seg]);
    if ((int32_t)(cur >> 32) < 0 || (uint32_t)(cur & 0xFFFF) != t || seg_home(cur) != p)
      return kNullBlock;
  }
  const uint64_t old = atomic_sub(&s->seg_state[seg], kActiveOne);
  const int32_t ac = (int32_t)(old >> 32);
  if (ac < 0 || (uint32_t)(old & 0xFFFF) != t || seg_home(old) != p) {
    atomic_add(&s->seg_state[seg], kActiveOne);
    return kNullBlock;
  }
  const uint64_t nb = nb_of(s, t);
  const uint32_t q = atomic_add(&s->seg_queue[seg].q_malloc, 1u);
  uint32_t blk;
  if ((uint64_t)q < nb) {
    blk = (uint32_t)((seg << s->bps_shift) + q);
  } else {
    blk = atomic_exch(&s->ring[(seg << s->bps_shift) + (q & (uint32_t)(nb - 1))],
                      kNullBlock);
  }
  if (blk == kNullBlock) {  // memory_table.cuh:977-982 traps here
    add_error(s, kStatErrRingEmpty);
    return kNullBlock;
  }
  // init_malloc (block.cuh:215-224): the tag carries class and home
  atomic_exch(&block_at(s, blk)->malloc_counter, block_tag(t, p) << kTreeShift);
  return (uint64_t)blk | ((uint64_t)(ac == 0) << 32);
}

// gallatin.cuh:1283-1302 submit_segment_for_free: the chunk id is reset
// before the run goes back to the segment tree, so a concurrent gather can
// never find the head still marked.
__device__ __noinline__ void free_segments(gal_state* s, uint64_t seg, uint64_t n) {
  clear_chunk_id(s, seg);
  utils::fence_release();
  veb::return_multiple(s->seg_tree, seg, n);
  add_stat(s, kStatSegmentsReleased, n);
}

// shared_block_storage.cuh:119-134 get_valid_block_index + get_my_block for
// the slots of class t in partition p: from the start slot skip slots
// without a block (at most 30 steps).  Returns (slot << 32) | block; with no
// block found, the start slot and its content (kNullBlock or kRefilling).
__device__ inline uint64_t pinned_block(const gal_state* s, uint32_t t, uint32_t p) {
  const uint32_t per = s->slot_count[t] / s->num_parts;
  const uint32_t base = p * per;
  uint32_t rel = 0;
  if (!deterministic(s)) {
    const uint32_t mix = ((__smid() & 63u) * 0x9E3779B1u) >> 26;  // 0..63
    rel = mix % per;
  }
  const uint32_t* sl = s->slots + s->slot_off[t] + base;
  const int64_t original = rel;
  int64_t i = rel;
  uint32_t blk = atomic_load(&sl[i]);
  const uint32_t first = blk;
  int counter = 0;
  while (blk >= kRefilling && i != original - 1) {
    i = (i + 1) % per;
    blk = atomic_load(&sl[i]);
    if (++counter >= kSlotScanCutoff) break;
  }
  if (blk >= kRefilling) {
    i = original;
    blk = first;
  }
  return ((uint64_t)(base + (uint32_t)i) << 32) | blk;
}

// --------------------------------------------------------- wave ops
// Functions below are called with identical arguments on every active lane
// and return wave-uniform results.

__device__ __forceinline__ bool try_lock(gal_state* s, uint32_t idx) {
  uint32_t got = 0;
  if (wave::is_first()) got = atomic_exch(&s->locks[idx], 1u) == 0u;
  return wave::readfirst(got) != 0;
}
__device__ __forceinline__ void unlock(gal_state* s, uint32_t idx) {
  utils::fence_release();
  if (wave::is_first()) atomic_store(&s->locks[idx], 0u);
}
__device__ __forceinline__ uint32_t tree_lock(const gal_state* s, uint32_t t, uint32_t p) {
  return 1 + t * s->num_parts + p;  // lock 0 guards multi-segment gathers
}

// memory_table.cuh:746-818 setup_segment: claim the chunk id for class t,
// home p, clear the ring (whole wave), zero the queue counters, then publish
// nb-1 free blocks.
__device__ inline bool setup_segment(gal_state* s, uint64_t seg, uint32_t t, uint32_t p) {
  uint32_t ok = 0;
  if (wave::is_first()) {
    uint64_t cur = atomic_load(&s->seg_state[seg]);
    for (int g = 0; g < 4096; ++g) {
      if ((cur & 0xFFFF) != kFreeTree) break;
      const uint64_t want = (cur & ~kSegIdMask) | seg_id(t, p);
      const uint64_t seen = atomic_cas(&s->seg_state[seg], cur, want);
      if (seen == cur) {
        ok = 1;
        break;
      }
      cur = seen;
    }
  }
  if (!wave::readfirst(ok)) return false;
  const uint64_t nb = nb_of(s, t);
  const uint64_t act = wave::active();
  const uint32_t k = wave::count(act);
  const uint32_t rank = wave::count(act & wave::lanes_below());
  for (uint64_t i = rank; i < nb; i += k)
    atomic_store(&s->ring[(seg << s->bps_shift) + i], kNullBlock);
  if (wave::is_first()) {
    gal_seg_queue* q = &s->seg_queue[seg];
    atomic_store(&q->q_malloc, 0u);
    atomic_store(&q->q_free, 0u);
    // a new generation, frontier 0, nothing marked
    const uint64_t gen = (atomic_load(&q->q_done) >> kDoneGenShift) + 1;
    atomic_store(&q->q_done, gen << kDoneGenShift);
  }
  release_to_xcd();
  if (wave::is_first()) {
    atomic_add(&s->seg_state[seg], nb << 32);  // active -1 -> nb-1
    add_stat(s, kStatSegmentsSetUp, 1);
  }
  return true;
}

// gallatin.cuh:1465-1510 gather_new_segment: smallest free segment.
// kNoSegment: the segment tree had none; kFail: set-up failed.
constexpr uint64_t kNoSegment = kFail - 1;
__device__ inline uint64_t gather_new_segment(gal_state* s, uint32_t t, uint32_t p) {
  const uint64_t seg = veb::take_first(s->seg_tree);
  if (seg == kFail) return kNoSegment;
  if (!setup_segment(s, seg, t, p)) {
    if (wave::is_first()) add_error(s, kStatErrSetup);
    return kFail;
  }
  if (wave::is_first()) veb::insert(class_tree(s, t, p), seg);
  return seg;
}

// Segment of class t (partition p) to take a block from: smallest id in
// deterministic mode (find_first_valid_index, veb.cuh:940-944), else from a
// hashed start (find_random_valid_index, veb.cuh:946-977) wrapping to 0.
__device__ inline uint64_t find_segment(gal_state* s, uint32_t t, uint32_t p, bool det) {
  const gal_tree& tr = class_tree(s, t, p);
  if (det) return veb::find_ge(tr, 0);
  const uint64_t gid = wave::readfirst(
      (uint64_t)threadIdx.x + (uint64_t)blockIdx.x * blockDim.x);
  const uint64_t start =
      hashers::MurmurHash64A_u64(gid, s->seed + t + 1) % tr.universe;
  uint64_t x = veb::find_ge(tr, start);
  if (x == kFail && start != 0) x = veb::find_ge(tr, 0);
  return x;
}

// gallatin.cuh:1525-1631 request_new_block_from_tree.  With no segment in
// the class tree and none free it polls (up to `polls` times, lim_request
// for a first request): frees in flight on other waves return segments
// within microseconds.
__device__ inline uint32_t request_new_block(gal_state* s, uint32_t t, uint32_t p,
                                             bool det, uint32_t polls) {
  int attempts = 0;
  uint64_t released = 0;  // segments given back so far, at the last check
  const uint32_t lk = tree_lock(s, t, p);
  for (uint32_t spins = 0; attempts < (int)polls && spins < kSpinCap;
       ++spins) {
    uint64_t seg = find_segment(s, t, p, det);
    if (seg == kFail) {
      if (veb::looks_empty(s->seg_tree)) {
        // nothing free: poll without taking the class lock, whose release
        // writes back the L2 and which every waiting wave would queue on.
        // Every 128 polls, give up when no segment came back in between:
        // then no free is in flight to wait for (blocks coming back to a
        // class tree are seen by the search itself).
        if ((attempts & 127) == 0) {
          const uint64_t rel =
              wave::readfirst((uint64_t)atomic_load(&s->stats[kStatSegmentsReleased]));
          if (attempts > 0 && rel == released) break;
          released = rel;
        }
        ++attempts;
        utils::backoff();
        continue;
      }
      if (!try_lock(s, lk)) {
        utils::backoff();
        continue;
      }
      seg = gather_new_segment(s, t, p);
      unlock(s, lk);
      if (seg == kFail || seg == kNoSegment) {
        ++attempts;
        if (seg == kNoSegment) utils::backoff();
        continue;
      }
    }
    uint64_t got = kNullBlock;
    if (wave::is_first()) got = get_block(s, seg, t, p, spins > 0);
    got = wave::readfirst(got);
    if (got >> 32) {
      // segment fully handed out: detach it and prefetch the next one
      // (gallatin.cuh:1611-1614).  With XCD partitions a segment whose
      // last block came back on a foreign XCD stays formatted in its home's
      // class tree, so the prefetch waits until that tree is empty: taking a
      // fresh segment for every reused one would format segments without
      // bound (a 1-block class does it on every block).
      if (wave::is_first()) veb::remove(class_tree(s, t, p), seg);
      if ((s->num_parts == 1 || veb::looks_empty(class_tree(s, t, p))) && try_lock(s, lk)) {
        gather_new_segment(s, t, p);
        unlock(s, lk);
      }
    }
    const uint32_t blk = (uint32_t)got;
    if (blk != kNullBlock) return blk;
  }
  return kNullBlock;
}

// Refill an empty pinned slot of class t (wave op).  Returns the block now
// in the slot; kRefilling when another wave is fetching it (try again
// later); kNullBlock when this wave fetched and none could be had (the slot
// is left empty for a later try).  One wave at a time fetches for a slot:
// it marks the slot kRefilling first.  The reference has no refill: a slot
// stays null once a replacement failed (gallatin.cuh:1091-1116).
__device__ __noinline__ uint32_t refill_slot(gal_state* s, uint32_t t, uint32_t p,
                                             uint32_t slot, bool det) {
  uint32_t* sp = s->slots + s->slot_off[t] + slot;
  uint32_t cur = kNullBlock;
  if (wave::is_first()) cur = atomic_cas(sp, kNullBlock, kRefilling);
  cur = wave::readfirst(cur);
  if (cur != kNullBlock) return cur;  // a block, or another wave's refill
  const uint32_t nb = request_new_block(s, t, p, det, s->lim_request);
  if (wave::is_first()) {
    atomic_exch(sp, nb);  // the block, or empty again
    if (nb != kNullBlock) add_stat(s, kStatRefills, 1);
  }
  return nb;
}

// gallatin.cuh:1083-1138 replace_block.  The detached slot holds kRefilling
// until the new block is in, so waiting waves do not fetch blocks of their
// own; if none can be had the slot is left empty.
__device__ __noinline__ void replace_block(gal_state* s, uint32_t t, uint32_t p,
                                           uint32_t slot, uint32_t blk, bool det) {
  uint32_t* sp = s->slots + s->slot_off[t] + slot;
  uint32_t seen = 0;
  if (wave::is_first()) {
    seen = atomic_cas(sp, blk, kRefilling);
    // the detached block normally has every slice claimed; if a delayed
    // helper detached a later incarnation that still has free slices,
    // retire those so the block can complete
    if (seen == blk &&
        (atomic_load(&block_at(s, blk)->malloc_counter) & kCountMask) < kSlicesPerBlock)
      retire_block(s, blk, p);
  }
  if (wave::readfirst(seen) != blk) return;
  const uint32_t nb = request_new_block(s, t, p, det, s->lim_request);
  if (wave::is_first()) atomic_exch(sp, nb);
}

// Slice phase of malloc: gallatin.cuh:865-1078 malloc_slice_allocation for
// every size class present in the wave at once.
__device__ __forceinline__ void malloc_slices(gal_state* s, bool want, uint32_t t,
                                              uint32_t cnt, uint32_t p,
                                              uint64_t& out_off) {
  const bool det = deterministic(s);
  const uint32_t lane = wave::lane();
  bool pending = want;
  uint32_t attempts = 0;
  for (;;) {
    const uint64_t pend = __ballot(pending);
    if (!pend) break;
    // coalesced groups by class (labeled_partition, gallatin.cuh:883-885)
    const uint64_t grp = wave::match(pend, pending, t);
    const int leader = pending ? wave::first(grp) : (int)lane;
    const bool is_leader = pending && leader == (int)lane;
    uint64_t sb = kNullBlock;
    if (is_leader) sb = pinned_block(s, t, p);
    // a group that finds no pinned block refills its start slot, groups in
    // lane order of their leaders
    for (uint64_t m = __ballot(is_leader && (uint32_t)sb == kNullBlock); m; m &= m - 1) {
      const int l = wave::first(m);
      const uint32_t sl = (uint32_t)(wave::readlane(sb, l) >> 32);
      const uint32_t got = refill_slot(s, wave::readlane(t, l), p, sl, det);
      if ((int)lane == l) sb = ((uint64_t)sl << 32) | got;
    }
    sb = wave::fetch(sb, leader);
    const uint32_t slot = (uint32_t)(sb >> 32), blk = (uint32_t)sb;
    const bool have = pending && blk < kRefilling;
    const bool waiting = pending && blk == kRefilling;  // another wave refills
    // exclusive prefix in lane order and group total (block.cuh:245-264)
    uint32_t prefix, total;
    if (__ballot(pending && cnt != 1) == 0) {
      prefix = wave::count(grp & wave::lanes_below());
      total = wave::count(grp);
    } else {
      wave::group_scan(pending ? cnt : 0u, grp, prefix, total);
    }
    uint32_t old = 0;
    if (have && is_leader) old = atomic_add(&block_at(s, blk)->malloc_counter, total);
    old = wave::fetch(old, leader);
    const uint32_t al = (old & kCountMask) + prefix;
    const bool start_valid = al <= 4095u;
    const bool end_valid = al + cnt <= 4096u;
    const bool should_replace = have && start_valid && (al + cnt > 4095u);
    uint32_t leftover = 0;
    if (have && start_valid) leftover = end_valid ? cnt - 1 : 4096u - al;
    // block_correct_frees (block.cuh:267-281); a credit that completes the
    // block returns it (the reference leaks such a block)
    if (__ballot(leftover != 0)) {
      uint32_t lp, lsum;
      wave::group_scan(leftover, have ? grp : 0ULL, lp, lsum);
      bool completes = false;
      if (have && is_leader && lsum) {
        const uint32_t of = atomic_add(&block_at(s, blk)->free_counter, lsum);
        completes = of + lsum == kSlicesPerBlock;
      }
      for (uint64_t m = __ballot(completes); m; m &= m - 1) {
        const uint32_t b = wave::readlane(blk, wave::first(m));
        if (wave::is_first()) {
          atomic_exch(&block_at(s, b)->free_counter, 0u);
          free_block(s, b, p);
        }
      }
    }
    // the group whose claim covers slice 4095 replaces the block; so does a
    // group that keeps finding its own class's block full (the claimer of
    // slice 4095 may have reached it through a stale slot read and detached
    // nothing); the slot CAS lets exactly one of them through
    const bool full_seen = have && !start_valid && attempts >= kHelpAfter &&
                           (old >> kTreeShift) == block_tag(t, p);
    const uint64_t rep = __ballot(should_replace || full_seen);
    for (uint64_t m = __ballot(is_leader && (grp & rep) != 0); m; m &= m - 1) {
      const int l = wave::first(m);
      replace_block(s, wave::readlane(t, l), p, wave::readlane(slot, l),
                    wave::readlane(blk, l), det);
    }
    // check_valid (block.cuh:343-357): the block still belongs to class t
    // (and to this partition)
    const bool valid = have && end_valid;
    const bool ok = valid && (old >> kTreeShift) == block_tag(t, p);
    if (ok) {
      out_off = (uint64_t)al + ((uint64_t)blk << 12);
      pending = false;
    }
    for (uint64_t m = __ballot(valid && !ok); m; m &= m - 1) {
      const uint32_t b = wave::readlane(blk, wave::first(m));
      if (wave::is_first()) block_release(s, b, 1, p);
    }
    // no block to be had: fail now rather than retry; a group waiting for
    // another wave's refill backs off and tries again
    const uint64_t nob = __ballot(pending && !have && !waiting);
    if (pending && !have && !waiting) pending = false;
    if (__ballot(waiting)) utils::backoff();
    const bool spent = pending && ++attempts >= s->lim_slice;
    if (spent) pending = false;
    const uint64_t spent_m = __ballot(spent);
    if ((nob | spent_m) && wave::is_first()) {
      if (nob) add_stat(s, kStatMissNoBlock, wave::count(nob));
      if (spent_m) add_stat(s, kStatMissRetries, wave::count(spent_m));
    }
  }
}

// gallatin.cuh:828-861 malloc_block_allocation + block.cuh:288-340
// malloc_fill_block, retried as gallatin.cuh:1226-1231 while blocks are
// handed out but lost to a concurrent claim.  Byte offset.
__device__ __noinline__ uint64_t malloc_block_path(gal_state* s, uint32_t bt, uint32_t p,
                                                   bool det, uint32_t polls) {
  for (uint32_t attempt = 0; attempt < s->lim_block; ++attempt) {
    const uint32_t blk = request_new_block(s, bt, p, det, polls);
    if (blk == kNullBlock) return kFail;
    uint32_t res = 0;
    if (wave::is_first()) {
      const uint32_t merged = atomic_add(&block_at(s, blk)->malloc_counter, 4096u);
      const uint32_t oldc = merged & kCountMask;
      if ((merged >> kTreeShift) != block_tag(bt, p)) {  // check_valid -> trap
        add_error(s, kStatErrTag);
      } else {
        atomic_add(&block_at(s, blk)->free_counter, 4095u - oldc);
        if (oldc != 0)
          block_release(s, blk, 1, p);
        else
          res = 1;
      }
    }
    if (wave::readfirst(res)) return offset_to_bytes(s, (uint64_t)blk << 12, bt);
  }
  return kFail;
}

// set_tree_id (memory_table.cuh:823-828) for the head of an n-segment run:
// chunk id T + 1 + n (per lane).  On failure the run goes back to the
// segment tree and the error is counted.
__device__ inline bool mark_multi_head(gal_state* s, uint64_t idx, uint64_t n) {
  const uint32_t T = s->num_trees;
  uint64_t cur = atomic_load(&s->seg_state[idx]);
  for (int g = 0; g < 4096; ++g) {
    if ((cur & 0xFFFF) != kFreeTree) break;
    const uint64_t seen = atomic_cas(&s->seg_state[idx], cur,
                                     (uint64_t)((cur & ~kSegIdMask) | (T + 1 + n)));
    if (seen == cur) return true;
    cur = seen;
  }
  add_error(s, kStatErrMultiHead);
  veb::return_multiple(s->seg_tree, idx, n);
  add_stat(s, kStatSegmentsReleased, n);  // segments back in the tree (see no_run_known)
  return false;
}

// A search for a run of n free segments that found none stays true until a
// segment goes back into the segment tree, and every path that puts one
// back adds to stat [3] after doing so.  So the smallest length a search
// found no run for, with stat [3] as read before that search, lets later
// requests of that length or more fail without taking the lock and scanning
// the tree again while stat [3] is unchanged.  (Between a return and its
// count a request can fail that would have found the returned segment; a
// search racing that return can miss it as well.)  Not used in the
// deterministic mode, whose replay scans as the reference does.
// Both are wave-uniform: every lane of the wave takes part in the search.
__device__ inline uint64_t released_count(const gal_state* s) {
  return wave::readfirst((uint64_t)atomic_load(&s->stats[kStatSegmentsReleased]));
}
__device__ inline bool no_run_known(const gal_state* s, uint64_t rel, uint64_t n) {
  const uint64_t h = wave::readfirst((uint64_t)atomic_load(&s->stats[kStatSegHint]));
  return h != 0 && (h >> 16) == (rel & 0xFFFFFFFFFFFFULL) && n >= (h & 0xFFFF);
}
__device__ inline void note_no_run(gal_state* s, uint64_t rel, uint64_t n) {
  const uint64_t mine = ((rel & 0xFFFFFFFFFFFFULL) << 16) | (n & 0xFFFF);
  const uint64_t h = atomic_load(&s->stats[kStatSegHint]);
  if ((h >> 16) != (mine >> 16) || (h & 0xFFFF) > n)
    atomic_store(&s->stats[kStatSegHint], (unsigned long long)mine);
}

// Take the multi-segment lock (lock 0, gallatin.cuh:796).
__device__ inline bool lock_segments(gal_state* s) {
  for (uint32_t spin = 0; spin < kSpinCap; ++spin) {
    if (try_lock(s, 0)) return true;
    utils::backoff();
  }
  return false;
}

// gallatin.cuh:790-826 malloc_segment_allocation under the segment lock,
// retried as gallatin.cuh:1246-1251.  Byte offset.  The lock is taken by one
// lane of one request at a time, so lanes of a wave never wait on each other.
__device__ __noinline__ uint64_t malloc_segment_path(gal_state* s, uint64_t n) {
  const uint32_t T = s->num_trees;
  if (n == 0 || n > s->num_segments || n + T + 1 >= kFreeTree) return kFail;
  const bool det = deterministic(s);
  for (uint32_t attempt = 0; attempt < s->lim_segment; ++attempt) {
    const uint64_t rel = released_count(s);
    if (!det && no_run_known(s, rel, n)) return kFail;
    if (!lock_segments(s)) return kFail;
    const uint64_t idx = veb::gather_multiple(s->seg_tree, n);
    if (idx == kFail) {  // no free run of n segments now
      unlock(s, 0);
      if (!det && wave::is_first()) note_no_run(s, rel, n);
      return kFail;
    }
    uint32_t ok = 0;
    if (wave::is_first()) ok = mark_multi_head(s, idx, n);
    ok = wave::readfirst(ok);
    unlock(s, 0);
    if (ok) return idx << s->seg_shift;
  }
  return kFail;
}

// The multi-segment requests of a wave (lanes with `want`, n segments each).
// Several of them claim one run of their total length under one taking of
// the lock and split it in lane order, each lane marking the head of its own
// part, so a wave serialises one claim instead of one per lane; every part
// is an ordinary n-segment allocation (freed on its own).  Without a run of
// the total (or in the deterministic mode, whose replay takes runs one lane
// at a time as the reference does) the lanes claim one after another, and a
// size that just found no run fails the wave's later requests of that size
// or more without another scan.  Byte offset per lane, kFail on a miss.
__device__ __noinline__ uint64_t malloc_segments(gal_state* s, bool want, uint64_t n,
                                                 bool det) {
  const uint32_t T = s->num_trees;
  const uint64_t wm = __ballot(want);
  const bool ok_n = n > 0 && n <= s->num_segments && n + T + 1 < kFreeTree;
  uint64_t out = kFail;
  if (!det && wave::count(wm) > 1 && !__ballot(want && !ok_n)) {
    uint32_t pre = 0, total = 0;
    wave::group_scan(want ? (uint32_t)n : 0u, wm, pre, total);
    const uint64_t rel = released_count(s);
    const bool hinted = total <= 0xFFFF;  // the hint holds 16-bit lengths
    if (total <= s->num_segments && !(hinted && no_run_known(s, rel, total)) &&
        lock_segments(s)) {
      const uint64_t idx = veb::gather_multiple(s->seg_tree, total);
      unlock(s, 0);
      if (idx != kFail) {
        if (want && mark_multi_head(s, idx + pre, n)) out = (idx + pre) << s->seg_shift;
        return out;
      }
      if (hinted && wave::is_first()) note_no_run(s, rel, total);
    }
  }
  uint64_t no_run = ~0ULL;  // smallest size that found no run
  for (uint64_t m = wm; m; m &= m - 1) {
    const int l = wave::first(m);
    const uint64_t nl = wave::readlane(n, l);
    const uint64_t r = nl >= no_run ? kFail : malloc_segment_path(s, nl);
    if (r == kFail && nl < no_run) no_run = nl;
    if ((int)wave::lane() == l) out = r;
  }
  return out;
}

// gallatin.cuh:1186-1276 malloc.  Lanes with want = false take part in the
// wave's cooperative work (tree searches) and get nullptr; every active lane
// of the wave must call it.
__device__ __forceinline__ void* malloc_impl(gal_state* s, uint64_t size, bool want = true) {
  const bool det = deterministic(s);
  const uint32_t T = s->num_trees;
  const uint32_t lane = wave::lane();
  const uint32_t p = wave_part(s);
  // size dispatch (gallatin.cuh:1141-1147, 1196-1258)
  const int cl = utils::get_first_bit_bigger(size < 1 ? 1 : size);
  uint32_t t = size < s->smallest ? 0u : (uint32_t)(cl - (int)s->smallest_bits);
  uint32_t path = want ? kPathSlice : 0u, cnt = 1;
  uint64_t nseg = 0;
  if (want && t >= T) {
    const int bt = cl - (int)s->block_tree_bits;
    if (bt < 0) {
      cnt = 1u << (t - (T - 1));
      t = T - 1;
    } else if (bt < (int)T) {
      path = kPathBlock;
      t = (uint32_t)bt;
    } else {
      path = kPathSegment;
      nseg = (size - 1) / s->bytes_per_segment + 1;
      t = 0;
    }
  }
  uint64_t bytes = kFail;
  if (__ballot(path == kPathSlice)) {
    uint64_t off = kFail;
    malloc_slices(s, path == kPathSlice, t, cnt, p, off);
    if (off != kFail) bytes = offset_to_bytes(s, off, t);
  }
  // The block requests of a wave run one after another.  Once one of them
  // has polled out its attempts with no segment to be had, the rest try
  // once each (any class tree may still hold a segment with a free block)
  // instead of polling the same empty pool again.
  uint32_t polls = s->lim_request;
  for (uint64_t m = __ballot(path == kPathBlock); m; m &= m - 1) {
    const int l = wave::first(m);
    const uint64_t r = malloc_block_path(s, wave::readlane(t, l), p, det, polls);
    if (r == kFail) polls = 1;
    if ((int)lane == l) bytes = r;
  }
  if (__ballot(path == kPathSegment)) {
    const uint64_t r = malloc_segments(s, path == kPathSegment, nseg, det);
    if (path == kPathSegment) bytes = r;
  }
  // No acquire on the way out.  The previous user of a slice released its
  // writes before its free (free_impl writes its XCD's L2 back, or the
  // slice stayed on its home XCD), so none of its bytes can land on the new
  // user's data later.  This XCD's L2 may still hold clean old copies of
  // the slice's lines; the L2 writes back only the bytes written, so those
  // copies show only in bytes the new user reads before writing them, which
  // malloc leaves undefined.  The allocator's own words are all accessed
  // with device-coherent atomics.  (An agent-scope acquire here invalidated
  // the XCD's L2 on every malloc of a one-partition instance.)
  return bytes == kFail ? nullptr : (void*)(s->memory + bytes);
}

// gallatin.cuh:1378-1461 free + 1774-1827 free_offset.
__device__ __forceinline__ void free_impl(gal_state* s, void* ptr) {
  const bool det = deterministic(s);
  const uint32_t T = s->num_trees;
  const uint32_t lane = wave::lane();
  const uint32_t my_p = wave_part(s);
  const uint64_t base = (uint64_t)s->memory;
  const uint64_t p = (uint64_t)ptr;
  const bool in = ptr != nullptr && (p - base) < (s->num_segments << s->seg_shift);
  uint64_t seg = 0, st = 0;
  uint32_t t = kFreeTree;
  if (in) {
    seg = (p - base) >> s->seg_shift;
    st = atomic_load(&s->seg_state[seg]);
    t = (uint32_t)(st & 0xFFFF);
  }
  const bool multi = in && t > T && t != kFreeTree;
  const bool slice = in && t < T;
  const uint64_t bad = __ballot(ptr != nullptr && !multi && !slice);
  if (bad && wave::is_first()) add_error(s, kStatErrInvalidFree, wave::count(bad));
  if (!__ballot(multi || slice)) return;
  if (!(s->flags & kFlagNoPayloadFence)) {
    // a free away from the segment's home XCD writes this XCD's L2 back
    if (s->num_parts == 1 || __ballot(multi || (slice && seg_home(st) != my_p)))
      utils::fence_release();
    else
      release_to_xcd();
  }
  // multi-segment frees one at a time in lane order
  for (uint64_t m = __ballot(multi); m; m &= m - 1) {
    const int l = wave::first(m);
    const uint64_t sg = wave::readlane(seg, l);
    const uint32_t tl = wave::readlane(t, l);
    if (wave::is_first()) free_segments(s, sg, tl - T - 1);
  }
  const uint64_t sm = __ballot(slice);
  if (!sm) return;
  uint32_t blk = 0;
  if (slice) blk = (uint32_t)(bytes_to_offset(s, p - base, t) >> 12);
  // group lanes freeing into the same block; in deterministic mode all of
  // them, else the first few groups and single-lane atomics for the rest
  uint64_t grp = 0, rem = sm;
  bool grouped = false;
  const int rounds = det ? 64 : 6;
  for (int r = 0; rem && r < rounds; ++r) {
    const uint32_t bl = wave::readlane(blk, wave::first(rem));
    const bool eq = slice && !grouped && blk == bl;
    const uint64_t g = __ballot(eq);
    if (eq) {
      grp = g;
      grouped = true;
    }
    rem &= ~g;
  }
  if (slice && !grouped) grp = 1ULL << lane;
  const bool is_leader = slice && wave::first(grp) == (int)lane;
  bool completes = false;
  if (is_leader) {
    const uint32_t k = wave::count(grp);
    const uint32_t old = atomic_add(&block_at(s, blk)->free_counter, k);
    completes = old + k == kSlicesPerBlock;
  }
  for (uint64_t m = __ballot(completes); m; m &= m - 1) {
    const uint32_t b = wave::readlane(blk, wave::first(m));
    if (wave::is_first()) {
      atomic_exch(&block_at(s, b)->free_counter, 0u);
      free_block(s, b, my_p);
    }
  }
}

// gallatin.cuh:1152-1182 get_allocated_size.
__device__ inline uint64_t allocated_size(const gal_state* s, uint64_t bytes) {
  if (bytes < 16) bytes = 16;
  const int bit_size = utils::get_first_bit_bigger(bytes);
  const uint32_t t = bytes < s->smallest ? 0u : (uint32_t)(bit_size - (int)s->smallest_bits);
  if (t > s->num_trees) {
    const int block_tree = bit_size - (int)s->block_tree_bits;
    if (block_tree > (int)s->num_trees) {
      const uint64_t n = (bytes - 1) / s->bytes_per_segment + 1;
      return n * s->bytes_per_segment;
    }
  }
  return 1ULL << bit_size;
}

}  // namespace detail

// Block counters under the reference's name (block.cuh:59-62: malloc_counter,
// free_counter).
using Block = gal_block;

// The reference's table accessors that its tests call through
// `allocator->table->` (tests/src/global_test.cu:105-107,
// global_test_noprint.cu:125-157).  `table` is the instance itself seen
// through this type, which adds no data.
struct memory_table : gal_state {
  // memory_table.cuh:1023-1027
  __device__ uint64_t get_segment_from_ptr(void* allocation) const {
    return ((uint64_t)allocation - (uint64_t)memory) >> seg_shift;
  }
  // memory_table.cuh:831-844: the chunk id, 0xFFFF for a free segment
  __device__ uint16_t read_tree_id(uint64_t segment) const {
    return (uint16_t)(utils::atomic_load(&seg_state[segment]) & 0xFFFFu);
  }
  // memory_table.cuh get_block_from_global_block_id: the counters of block id
  __device__ Block* get_block_from_global_block_id(uint64_t global_block_id) const {
    return blocks + block_position(global_block_id, block_bits);
  }
  // memory_table.cuh:1057-1061
  __device__ uint64_t get_blocks_per_segment(uint16_t tree) const {
    return blocks_per_segment >> tree;
  }
  // memory_table.cuh:1090-1127: slice offset of an allocation of class tree
  __device__ uint64_t allocation_to_offset(void* allocation, uint16_t tree) const {
    return detail::bytes_to_offset(this, (uint64_t)allocation - (uint64_t)memory, tree);
  }
  // memory_table.cuh:1346-1353
  __device__ bool owns_allocation(void* allocation) const {
    return ((uint64_t)allocation - (uint64_t)memory) < (num_segments << seg_shift);
  }
};

// Drop-in for gallatin::allocators::Gallatin (gallatin.cuh:292-2035).
template <uint64_t SegBytes, uint64_t Smallest, uint64_t Biggest>
struct Gallatin : public gal_state {
  using my_type = Gallatin<SegBytes, Smallest, Biggest>;
  static_assert(SegBytes >= Biggest * 4096,
                "a segment must hold one block of the biggest class");
  static_assert((SegBytes & (SegBytes - 1)) == 0 && (Smallest & (Smallest - 1)) == 0 &&
                    (Biggest & (Biggest - 1)) == 0 && Smallest <= Biggest,
                "sizes must be powers of two");

  // gallatin.cuh:324-421.  Returns nullptr (after printing why) on failure;
  // calloc mode is not supported.
  static __host__ my_type* generate_on_device(uint64_t max_bytes, uint64_t seed,
                                              bool print_info = true,
                                              bool running_calloc = false) {
    return generate_with_pool(max_bytes, seed, print_info, running_calloc, 0u);
  }

  // gallatin.cuh:427-533: the pool in pinned, mapped host memory; the
  // metadata stays in device memory.
  static __host__ my_type* generate_on_device_host(uint64_t max_bytes, uint64_t seed,
                                                   bool print_info = true,
                                                   bool running_calloc = false) {
    return generate_with_pool(max_bytes, seed, print_info, running_calloc, GAL_FLAG_POOL_HOST);
  }

  // gallatin.cuh:536-630: the pool in managed memory.
  static __host__ my_type* generate_on_device_managed(uint64_t max_bytes, uint64_t seed,
                                                      bool print_info = true,
                                                      bool running_calloc = false) {
    return generate_with_pool(max_bytes, seed, print_info, running_calloc,
                              GAL_FLAG_POOL_MANAGED);
  }

  static __host__ my_type* generate_with_pool(uint64_t max_bytes, uint64_t seed, bool print_info,
                                              bool running_calloc, uint32_t pool_flag) {
    if (running_calloc) {
      fprintf(stderr, "Gallatin: calloc mode is not supported\n");
      return nullptr;
    }
    gal_config cfg;
    gal_config_default(&cfg, max_bytes, seed);
    cfg.bytes_per_segment = SegBytes;
    cfg.smallest = Smallest;
    cfg.biggest = Biggest;
    cfg.print_info = print_info ? 1u : 0u;
    cfg.flags |= pool_flag;
    gal_allocator* h = nullptr;
    if (gal_create(&cfg, &h) != GAL_OK) {
      fprintf(stderr, "Gallatin: generate_on_device failed: %s\n", gal_last_error());
      return nullptr;
    }
    return static_cast<my_type*>(gal_device_state(h));
  }

  // gallatin.cuh:643-668.
  static __host__ void free_on_device(my_type* dev_version) {
    gal_destroy(gal_from_device_state(dev_version));
  }

  // gallatin.cuh:1855-1897 (called on the device pointer, as in the reference).
  __host__ void print_info() { gal_print_info(gal_from_device_state(this)); }

  static __host__ __device__ int get_first_bit_bigger(uint64_t x) {
    return utils::get_first_bit_bigger(x);
  }
  static __host__ __device__ int get_num_trees() {
    return get_first_bit_bigger(Biggest) - get_first_bit_bigger(Smallest) + 1;
  }
  static __host__ __device__ uint64_t get_blocks_per_segment(uint16_t tree) {
    return SegBytes / ((Smallest << tree) * 4096);
  }

  __device__ void* malloc(uint64_t size) { return detail::malloc_impl(this, size); }
  __device__ void free(void* ptr) { detail::free_impl(this, ptr); }
  __device__ uint64_t get_allocated_size(uint64_t bytes) {
    return detail::allocated_size(this, bytes);
  }
  // gallatin.cuh:2027 / memory_table.cuh:1346-1353
  __device__ bool owns_allocation(void* p) {
    return ((uint64_t)p - (uint64_t)memory) < (num_segments << seg_shift);
  }
  // gallatin.cuh:1842-1846
  __device__ uint64_t allocation_to_offset(void* allocation, uint16_t tree_id) {
    return table->allocation_to_offset(allocation, tree_id);
  }
  // gallatin.cuh:673-690: whether the allocation lies inside the pool, with
  // the reference's diagnostic lines
  __device__ bool check_alloc_valid(void* allocation) {
    const uint64_t byte_difference = (uint64_t)allocation - (uint64_t)memory;
    const uint64_t max_bytes = num_segments * bytes_per_segment;
    printf("%llx inside of %llx\n", (unsigned long long)allocation,
           (unsigned long long)memory);
    if (byte_difference > max_bytes) {
      printf("Byte difference: %llu > %llu\n", (unsigned long long)byte_difference,
             (unsigned long long)max_bytes);
      return false;
    }
    printf("Byte difference: %llu <= %llu\n", (unsigned long long)byte_difference,
           (unsigned long long)max_bytes);
    return true;
  }
};

}  // namespace allocators
}  // namespace gallatin

#endif

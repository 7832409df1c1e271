// HBM layout of one allocator instance.  Shared by the device code in
// gallatin.hpp and the host library that creates instances (libgallatin_hip).
//
// Reference structures this replaces:
//   veb_tree / layer        veb.cuh:109-1228   (one cudaMalloc per layer)
//   alloc_table             memory_table.cuh:229-286
//   Block                   block.cuh:59-398
//   pinned_shared_blocks    shared_block_storage.cuh:48-304
//
// Differences in layout (all integer, all agent-scope atomics):
//   * all tree levels of all trees live in one arena of 64-bit words; a level
//     of w words is scanned by the wave 64 words per pass;
//   * the per-segment chunk id (u16) and active count (int) are packed into
//     one 64-bit word `seg_state` = (active << 32) | tree_id, so `free` reads
//     one aligned word per pointer and `get_block` learns count and owner in
//     one atomic;
//   * ring entries and pinned slots hold 32-bit block ids instead of Block*.
#ifndef GALLATIN_LAYOUT_HPP
#define GALLATIN_LAYOUT_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gallatin {
namespace allocators {

constexpr uint32_t kSlicesPerBlock = 4096;   // block.cuh:84-97
constexpr uint32_t kTreeShift = 20;          // block.cuh:53
constexpr uint32_t kCountMask = (1u << kTreeShift) - 1u;
constexpr uint32_t kFreeTree = 0xFFFFu;      // memory_table.cuh:485
constexpr uint32_t kNullBlock = 0xFFFFFFFFu;
// Pinned-slot marker: a wave is fetching the slot's next block.  Other waves
// wait for it instead of fetching blocks of their own.
constexpr uint32_t kRefilling = 0xFFFFFFFEu;
constexpr uint64_t kFail = ~0ULL;
constexpr int kMaxLayers = 4;                // universes up to 2^24 segments
constexpr int kMaxTrees = 24;
constexpr int kMaxParts = 8;                 // XCD partitions

// retry limits (gallatin.cuh:108-112, shared_block_storage.cuh:31)
constexpr int kRequestBlockAttempts = 1000;
constexpr int kSliceAttempts = 500 * 10;
constexpr int kBlockAttempts = 500;
constexpr int kSegmentAttempts = 500;
constexpr int kSlotScanCutoff = 30;
// failed claims on a full block before a group helps to replace it
constexpr uint32_t kHelpAfter = 16;
// bound on loops the reference leaves unbounded (lost lock races)
constexpr uint32_t kSpinCap = 1u << 16;

// flags
constexpr uint32_t kFlagDeterministic = 1u;  // smallest-id segment, slot 0
constexpr uint32_t kFlagNoPayloadFence = 2u; // skip the payload release at free

// seg_state packing: bits 0-15 tree id (kFreeTree = unformatted), bits
// 16-23 home partition, bits 32-63 active count (int32)
constexpr uint64_t kActiveOne = 1ULL << 32;
constexpr uint64_t kSegIdMask = 0xFFFFFFULL;
constexpr uint64_t kSegFreeState = 0xFFFFFFFF0000FFFFULL;  // active -1, tree free

__host__ __device__ inline uint32_t seg_home(uint64_t st) { return (uint32_t)(st >> 16) & 0xFFu; }
__host__ __device__ inline uint64_t seg_id(uint32_t tree, uint32_t home) {
  return (uint64_t)tree | ((uint64_t)home << 16);
}
// Tag stored above the count of a block's malloc_counter: class and home
// partition (just the class with one partition, as in block.cuh:215-224).
__host__ __device__ inline uint32_t block_tag(uint32_t tree, uint32_t home) {
  return tree | (home << 5);
}

struct gal_block {
  uint32_t malloc_counter;  // (tree << 20) | claimed slices
  uint32_t free_counter;    // freed + pre-credited slices
};

// Position of block b's counters in the block array of 2^k entries.  The
// blocks that are hot at the same time (pinned blocks, numbered segment by
// segment) would otherwise share a few 4 KiB regions, and device-scope
// atomics on one region serialise at one memory-side unit (~90 M/s); this
// bijection of [0, 2^k) (xorshift, odd multiply, xorshift) scatters
// neighbouring ids over the whole array.
__host__ __device__ inline uint64_t block_position(uint64_t b, uint32_t k) {
  if (k < 10) return b;
  const uint32_t h = (k + 1) / 2;
  const uint64_t mask = (1ULL << k) - 1;
  uint64_t x = b ^ (b >> h);
  x = (x * 0x9E3779B97F4A7C15ULL) & mask;
  return x ^ (x >> h);
}

struct gal_tree {
  uint64_t* words;                     // layer 0 (top) .. num_layers-1 (bottom)
  uint64_t universe;                   // bits in the bottom layer
  uint32_t num_layers;
  uint32_t pad;
  uint64_t layer_off[kMaxLayers];      // word offset of each layer in `words`
  uint64_t layer_words[kMaxLayers];
};

struct gal_seg_queue {
  uint32_t q_malloc;  // ring read position   (memory_table.cuh:244)
  uint32_t q_free;    // ring write ticket    (memory_table.cuh:247)
  // ring write completion (memory_table.cuh:249): set-up generation in bits
  // 63..56, the first incomplete ticket mod 2^24 in bits 55..32, and in the
  // low half the written-mask of it and the next 31 tickets (ticket x at
  // bit x mod 32)
  uint64_t q_done;
};

// stats[] indices.  kStatErrors is the sum of the kinds from kStatErrFirst
// to kStatErrInvalidFree.
enum gal_stat {
  kStatErrors = 0,
  kStatBootFailed = 1,
  kStatSegmentsSetUp = 2,
  kStatSegmentsReleased = 3,
  kStatBlocksReturned = 4,
  kStatRefills = 5,         // empty pinned slots given a block again
  kStatErrFirst = 6,
  kStatErrChunkId = 6,      // a released segment's chunk id could not be reset
  kStatErrRingOrder = 7,    // ring completion never came round
  kStatErrBlockFree = 8,    // block returned into an unformatted segment
  kStatErrRingEmpty = 9,    // ring entry empty where a block was due
  kStatErrSetup = 10,       // a free segment could not be formatted
  kStatErrTag = 11,         // whole-block claim found another class's tag
  kStatErrMultiHead = 12,   // multi-segment head could not be marked
  kStatErrInvalidFree = 13, // nullptr, foreign or unformatted-segment frees
  kStatMissNoBlock = 14,    // slice requests failed: no block to be had
  kStatMissRetries = 15,    // slice requests failed: retries used up
  kStatRingRuns = 16,       // ring completions that covered several tickets
  kStatRingRunTickets = 17, // tickets those completions covered
  kStatCount = 18,
  // internal, after the reported counters: the last multi-segment search
  // that found no run, as (stat [3] when it started) << 16 | run length
  kStatSegHint = 18,
  kStatWords = 19
};

// The instance viewed through the reference's alloc_table accessors
// (memory_table.cuh:229-1353); defined in gallatin.hpp.
struct memory_table;

struct gal_state {
  // geometry, fixed at creation
  uint64_t bytes_per_segment;
  uint64_t smallest;
  uint64_t biggest;
  uint64_t num_segments;
  uint64_t blocks_per_segment;   // S / (smallest * 4096)
  uint64_t seed;
  uint32_t seg_shift;            // log2(bytes_per_segment)
  uint32_t smallest_bits;        // log2(smallest)
  uint32_t bps_shift;            // log2(blocks_per_segment)
  uint32_t num_trees;
  uint32_t block_tree_bits;      // log2(smallest * 4096)
  uint32_t flags;
  uint32_t total_slots;
  uint32_t block_bits;           // block array holds 2^block_bits entries
  uint32_t num_parts;            // XCD partitions (1 in deterministic mode)
  uint32_t pad1;
  // retry limits (defaults kRequestBlockAttempts, kSliceAttempts, ...)
  uint32_t lim_request;
  uint32_t lim_slice;
  uint32_t lim_block;
  uint32_t lim_segment;
  // storage
  char* memory;                  // num_segments * bytes_per_segment
  uint64_t* seg_state;           // [N]
  gal_seg_queue* seg_queue;      // [N]
  uint32_t* ring;                // [N * B] recycled block ids
  gal_block* blocks;             // [2^block_bits], at block_position()
  uint32_t* slots;               // pinned blocks of all size classes
  uint32_t* locks;               // [1 + T * parts]: multi-segment lock, then one
                                 // per class tree (gallatin.cuh:314, 1513-1519)
  unsigned long long* stats;     // [kStatWords]: kStatCount reported, then internal words
  uint32_t slot_off[kMaxTrees + 1];
  uint32_t slot_count[kMaxTrees];  // multiple of num_parts; partition p owns
                                   // slots [p, p+1) * count / num_parts
  gal_tree seg_tree;
  gal_tree sub_trees[kMaxTrees * kMaxParts];  // class t, partition p at t*parts+p
  // `table` of the reference's Gallatin (gallatin.cuh:306): the device
  // address of this instance, viewed as a memory_table.
  memory_table* table;
};

}  // namespace allocators
}  // namespace gallatin

#endif

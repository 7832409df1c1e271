/*
 * TEST INFRASTRUCTURE ONLY - the oracle.
 *
 * A sequential, single-threaded C restatement of the Gallatin allocator
 * policy (saltsystemslab/gallatin, include/gallatin/allocators/ headers).  It is
 * the checker for the HIP implementation: tests, the smoke entry point and
 * the CPU-baseline leg of the benchmark load it; no product code path links
 * or calls it.
 *
 * The reference is nondeterministic in three places (racy boot, %smid slot
 * choice, thread-id-hashed segment start).  The oracle fixes them as the
 * product's deterministic mode does:
 *   boot      = boot_shared_block_container_one_thread order (gallatin.cuh:195-221)
 *   segment   = smallest id in the size class tree: find_first_valid_index
 *               (veb.cuh:940-944)
 *   slot      = get_valid_block_index with smid 0 (shared_block_storage.cuh:119-134)
 * and defines the order of the work one 64-lane wave step performs (the
 * reference leaves it to cooperative-groups scheduling):
 *   malloc: slice lanes, then block lanes in lane order, then multi-segment
 *           lanes in lane order; slice lanes are grouped by size class, groups
 *           ordered by their lowest lane, lanes inside a group take
 *           consecutive slices in lane order (exclusive prefix sum).
 *   free:   multi-segment lanes in lane order, then slice/block lanes grouped
 *           by block, block returns in order of each group's lowest lane.
 *
 * Deviations from the reference (each is a reference defect; see the
 * allocator's design notes):
 *   F1 layer return sets a parent bit when the child word becomes non-empty
 *      (veb.cuh:1140 sets it only when the child becomes full).
 *   F2 multi-segment gather also considers word 0 and never walks below bit 0
 *      (veb.cuh:331 skips word 0; the inner walk can index word -1).
 *   F3 a multi-segment free resets the chunk id before returning the bits
 *      (gallatin.cuh:1285-1289 does it the other way round, which races).
 *   F4 the free-counter credit for a claim that straddles slice 4096 returns
 *      the block if the credit completes it (block.cuh:267-281 never checks,
 *      so such a block leaks).
 *   F5 free(nullptr), frees outside the pool and frees into an unformatted
 *      segment are counted as errors and ignored (reference: UB / trap).
 *   F6 a coalesced group that finds no pinned block of its class refills the
 *      empty start slot with a new block (the reference leaves a slot null
 *      once a replacement failed, gallatin.cuh:1091-1116, so a class that
 *      ran out of memory never allocates again); a whole-block request
 *      fails once no block can be had and a multi-segment request fails
 *      when no free run of n segments exists, instead of repeating the same
 *      search 500 times (gallatin.cuh:1226-1251).
 *   F7 the 10 % re-register of a segment into its class tree follows the
 *      segment's fill: a block return that brings its free blocks to at
 *      least ceil(nb / 10) without completing it inserts the segment when
 *      it is not in the tree.  The reference tests the ring position
 *      pos % nb of the returned block instead (gallatin.cuh:1646 with
 *      memory_table.cuh:1175-1182, marked "TODO - make this actually
 *      calculate fill"), so a segment whose last block was handed out at
 *      ring position k becomes findable again only when its ring wraps to
 *      10 %: on average half of its blocks free and unusable.
 */
#ifndef GALLATIN_ORACLE_H
#define GALLATIN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_FAIL (~0ULL)
#define ORC_MAX_LAYERS 8
#define ORC_MAX_TREES 24

/* ---------------------------------------------------------------- utils */
int orc_clog2(uint64_t x);
int orc_cfcll(uint64_t bits);
uint64_t orc_murmur64a(const void *key, int len, uint64_t seed);
void orc_chain_of(const void *keys, int key_bytes, uint64_t n, uint64_t seed,
                  uint64_t nchains, uint64_t *out);
void orc_big_stream(uint64_t seed, uint32_t step, uint64_t n_small, uint64_t n_large,
                    uint64_t *sizes, uint64_t *order);

/* ------------------------------------------------------------- vEB tree */
typedef struct orc_veb {
  uint64_t universe;
  int num_layers;
  uint64_t layer_universe[ORC_MAX_LAYERS]; /* index 0 = top layer */
  uint64_t layer_words[ORC_MAX_LAYERS];
  uint64_t *bits[ORC_MAX_LAYERS];
} orc_veb;

orc_veb *orc_veb_create(uint64_t universe, int filled);
void orc_veb_destroy(orc_veb *t);
int orc_veb_num_layers(const orc_veb *t);
uint64_t orc_veb_layer_words(const orc_veb *t, int layer);
uint64_t *orc_veb_layer_bits(orc_veb *t, int layer);
int orc_veb_insert(orc_veb *t, uint64_t x);  /* insert_force_update */
int orc_veb_remove(orc_veb *t, uint64_t x);
int orc_veb_query(const orc_veb *t, uint64_t x);
uint64_t orc_veb_successor(const orc_veb *t, uint64_t x); /* strictly greater */
uint64_t orc_veb_find_first(const orc_veb *t);
uint64_t orc_veb_malloc_first(orc_veb *t);
uint64_t orc_veb_gather_multiple(orc_veb *t, uint64_t n);
void orc_veb_return_multiple(orc_veb *t, uint64_t start, uint64_t n);
uint64_t orc_veb_fill(const orc_veb *t);

/* ------------------------------------------------------------ allocator */
typedef struct orc_alloc orc_alloc;

/* Returns NULL when the geometry is invalid or the boot runs out of
 * segments. slots0/slots_min: pinned-slot counts (reference 128 / 4). */
orc_alloc *orc_create(uint64_t max_bytes, uint64_t bytes_per_segment,
                      uint64_t smallest, uint64_t biggest, uint32_t slots0,
                      uint32_t slots_min);
void orc_destroy(orc_alloc *a);

/* Retry limits (<= 0 keeps the current value).  Defaults are the
 * reference's: request_new_block 1000, slice 500 x 10, block 500, segment
 * 500 (gallatin.cuh:108-112).  In a sequential run an attempt that fails
 * after the first retry changes no state, so results do not depend on them
 * as long as each is >= 3; they only bound the time spent out of memory. */
void orc_set_limits(orc_alloc *a, int request, int slice, int block,
                    int segment);

/* One wave step.  mask: active lanes; sizes/offsets indexed by lane.
 * Offsets are byte offsets from the pool base; ORC_FAIL = nullptr. */
void orc_malloc_wave(orc_alloc *a, uint64_t mask, const uint64_t *sizes,
                     uint64_t *out_offsets);
void orc_free_wave(orc_alloc *a, uint64_t mask, const uint64_t *offsets);

/* Convenience single-lane forms. */
uint64_t orc_malloc(orc_alloc *a, uint64_t size);
void orc_free(orc_alloc *a, uint64_t offset);

/* Wave-step stream replay.  Each step: kind (0 malloc, 1 free), mask and 64
 * arguments (malloc: size; free: index step*64+lane of an earlier malloc
 * result).  results[step*64+lane] receives malloc offsets (ORC_FAIL
 * otherwise). Returns 0, or -1 on a malformed stream. */
typedef struct orc_wave_step {
  uint32_t kind;
  uint32_t pad;
  uint64_t mask;
  uint64_t arg[64];
} orc_wave_step;
int orc_run_stream(orc_alloc *a, const orc_wave_step *steps, uint64_t nsteps,
                   uint64_t *results);

/* Return pinned blocks and untouched formatted segments (quiesced only). */
void orc_drain(orc_alloc *a);

/* Geometry / state inspection. */
typedef struct orc_geometry {
  uint64_t num_segments, bytes_per_segment, smallest, biggest;
  uint64_t blocks_per_segment;
  uint32_t num_trees, smallest_bits, total_slots, pad;
  uint64_t errors;
} orc_geometry;
void orc_get_geometry(const orc_alloc *a, orc_geometry *g);
orc_veb *orc_segment_tree(orc_alloc *a);
orc_veb *orc_sub_tree(orc_alloc *a, int tree);
/* Per segment: tree id (0xFFFF free) and active count. */
void orc_segment_state(const orc_alloc *a, uint16_t *tree_ids, int32_t *active);
/* Pinned slots: block id or 0xFFFFFFFF. */
void orc_slots(const orc_alloc *a, uint32_t *slots);
/* Block counters, 2 uint32 per block. */
void orc_blocks(const orc_alloc *a, uint32_t *counters);

/* Sequential per-thread churn (cpu baseline / consistency): threads
 * [tid0, tid0+nthreads) each run `rounds` rounds of the mixed-size churn of
 * streams.h one thread at a time.  Returns successful malloc+free pairs;
 * *misses counts nullptr results, *bad pattern mismatches. */
uint64_t orc_run_churn_sequential(orc_alloc *a, uint64_t seed, uint64_t tid0,
                                  uint64_t nthreads, uint32_t rounds,
                                  uint64_t *misses, uint64_t *bad);

#ifdef __cplusplus
}
#endif

#endif

/*
 * C ABI of libgallatin_hip.so: host lifecycle of device allocators and the
 * workload kernels that drive the device malloc/free path.
 *
 * The reference (saltsystemslab/gallatin) is header-only CUDA C++ with no FFI;
 * its host API is the static/host members of Gallatin<> and the free
 * functions of global_allocator.cuh.  Each entry point below names the
 * reference interface it replaces (paths relative to the reference checkout,
 * include/gallatin/allocators/ unless stated otherwise).  Device-side
 * malloc/free stay header-inline (include/gallatin/allocators/gallatin.hpp),
 * as in the reference, because device code cannot cross a C ABI.
 *
 * Conventions: every function returns GAL_OK (0) or a negative gal_status;
 * nothing calls exit().  Pointers named *_dev are device pointers; `stream`
 * is a hipStream_t (NULL = default stream).  Workload entry points enqueue
 * work on `stream` and return without synchronising unless stated.
 */
#ifndef GALLATIN_CAPI_H
#define GALLATIN_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gal_status {
  GAL_OK = 0,
  GAL_ERR_INVALID = -1,     /* bad argument or geometry */
  GAL_ERR_HIP = -2,         /* a HIP runtime call failed */
  GAL_ERR_BOOT = -3,        /* pool too small to pin the boot blocks */
  GAL_ERR_UNSUPPORTED = -4, /* e.g. calloc mode */
  GAL_ERR_NO_DEVICE = -5
} gal_status;

/* flags */
#define GAL_FLAG_DETERMINISTIC 1u     /* smallest-id segments, slot scan from 0 */
#define GAL_FLAG_NO_PAYLOAD_FENCE 2u  /* caller orders payload itself */
/* Pool placement (metadata stays in device memory either way):
 * GAL_FLAG_POOL_HOST: pinned, mapped host memory (generate_on_device_host,
 * gallatin.cuh:427-533; memory_table.cuh:521-542); GAL_FLAG_POOL_MANAGED:
 * managed memory (generate_on_device_managed, gallatin.cuh:536-630;
 * memory_table.cuh:543-560).  At most one of the two. */
#define GAL_FLAG_POOL_HOST 4u
#define GAL_FLAG_POOL_MANAGED 8u
/* No boot: the pinned slots start empty and the first request that finds a
 * slot empty fills it (the refill path), so segments are formatted only for
 * the classes and XCD partitions a workload uses.  This lets a small pool
 * run one partition per XCD (for example 1 GiB with one size class), where
 * the eager boot of every class in every partition would not fit.  Not
 * with GAL_FLAG_DETERMINISTIC, whose replay follows the eager boot order
 * (boot_shared_block_container, gallatin.cuh:195-221). */
#define GAL_FLAG_LAZY_BOOT 16u

typedef struct gal_config {
  uint64_t max_bytes;          /* pool bytes; segments = max_bytes / bytes_per_segment */
  uint64_t bytes_per_segment;  /* power of two, >= biggest * 4096 */
  uint64_t smallest;           /* smallest slice, power of two */
  uint64_t biggest;            /* biggest slice, power of two */
  uint64_t seed;
  uint32_t flags;
  uint32_t slots0;             /* pinned blocks of the smallest class (reference 128) */
  uint32_t slots_min;          /* floor of the halving slot counts (reference 4) */
  uint32_t print_info;         /* print the boot line, as generate_on_device does */
  /* retry limits, 0 = the reference's (gallatin.cuh:108-112): block
   * requests 1000, slice claims 500 x 10, whole blocks 500, segment runs 500 */
  uint32_t request_attempts;
  uint32_t slice_attempts;
  uint32_t block_attempts;
  uint32_t segment_attempts;
  /* XCD partitions of the pinned slots and class trees: 0 = one per XCD of
   * the device (8 on MI355X); forced to 1 in deterministic mode.  Slot
   * counts are rounded up to a multiple of it. */
  uint32_t partitions;
  uint32_t pad;
} gal_config;

typedef struct gal_allocator gal_allocator;

typedef struct gal_info {
  uint64_t num_segments, bytes_per_segment, smallest, biggest;
  uint64_t blocks_per_segment;
  uint32_t num_trees, total_slots;
  uint32_t partitions, pad;
  uint64_t segments_free;          /* bits set in the segment tree */
  uint64_t segments_formatted;     /* segments owned by a size class */
  uint64_t segments_multi;         /* heads of multi-segment allocations */
  uint64_t tree_owned[24];         /* bits set in each size-class tree */
  uint64_t errors;                 /* invalid frees / internal inconsistencies */
  uint64_t overhead_bytes;         /* metadata bytes (calculate_overhead) */
  uint64_t pool_base;              /* device address of the pool */
  /* print_guided_fill (gallatin.cuh:1958-1990) per size class t: segments
   * formatted for t and not closed, their blocks not yet handed out
   * (active count + 1 each), and their blocks handed out (the rest). */
  uint64_t class_segments[24];
  uint64_t class_blocks_free[24];
  uint64_t class_blocks_used[24];
} gal_info;

/* Fill `cfg` with the reference defaults: Gallatin<16 MiB, 16, 4096>,
 * 128/4 pinned slots (gallatin.cuh:346, 124; global_allocator.cuh:41). */
void gal_config_default(gal_config *cfg, uint64_t max_bytes, uint64_t seed);

/* Gallatin<S,m,M>::generate_on_device (gallatin.cuh:324-421): allocate the
 * pool and metadata on the current device, boot the pinned blocks and
 * return a handle.  Synchronous. */
int gal_create(const gal_config *cfg, gal_allocator **out);

/* Gallatin<S,m,M>::free_on_device (gallatin.cuh:643-668). */
int gal_destroy(gal_allocator *a);

/* Device address of the allocator object, usable as Gallatin<S,m,M>* in
 * kernels (the pointer generate_on_device returns). */
void *gal_device_state(gal_allocator *a);

/* Handle owning a device allocator object, or NULL (lets the C++ wrapper
 * serve print_info() called on a device pointer, gallatin.cuh:1855). */
gal_allocator *gal_from_device_state(const void *dev_state);

/* print_info (gallatin.cuh:1855-1897): segment-tree fill, per-class owned
 * segments, the per-class guided fill (print_usage, gallatin.cuh:1946-1990),
 * overhead.  Synchronous. */
int gal_print_info(gal_allocator *a);

/* Statistics behind print_info as numbers.  Synchronous. */
int gal_get_info(gal_allocator *a, gal_info *out);

/* Event counters of the instance (no counterpart in the reference; the
 * reference prints errors instead).  Copies min(n, GAL_STAT_COUNT) values
 * and returns GAL_STAT_COUNT.  [0] errors (sum of [6..13]), [1] boot
 * failed, [2] segments formatted, [3] segments given back to the free
 * pool (single segments whose last block came back, the segments of
 * multi-segment frees, and runs a failed multi-segment claim gave back),
 * [4] blocks
 * returned, [5] empty pinned slots refilled, [6] chunk id reset
 * failed, [7] ring completion timed out, [8] block freed into an
 * unformatted segment, [9] empty ring entry, [10] segment set-up failed,
 * [11] whole-block tag mismatch, [12] multi-segment head marking failed,
 * [13] invalid frees, [14] slice requests failed for want of a block, [15]
 * slice requests failed after all retries, [16] segment-ring completions
 * that covered more than one returned block, [17] the returned blocks they
 * covered.  Synchronous. */
#define GAL_STAT_COUNT 18
int gal_get_stats(gal_allocator *a, uint64_t *out, uint32_t n);

/* Quiescent drain (no counterpart in the reference): give back every pinned
 * block's unclaimed slices and release formatted segments whose blocks are
 * all free, so that after all frees the metadata equals a fresh instance.
 * Pinned slots are empty afterwards; call gal_reboot before new work. */
int gal_drain(gal_allocator *a, void *stream);

/* Re-pin boot blocks after a drain (boot_shared_block_container order); a
 * no-op for an instance created with GAL_FLAG_LAZY_BOOT. */
int gal_reboot(gal_allocator *a, void *stream);

/* Copy metadata to host buffers for checking (any pointer may be NULL):
 * seg_state[N] (active << 32 | tree id), slots[total_slots], blocks[2*N*B]
 * (malloc_counter, free_counter).  Synchronous. */
int gal_snapshot(gal_allocator *a, uint64_t *seg_state, uint32_t *slots,
                 uint32_t *blocks);

/* Tree words: tree -1 = segment tree, 0..T-1 = size-class trees (of XCD
 * partition `partition`).  Writes the words of all layers, top layer first,
 * and their counts.  Returns the number of words needed when out is NULL. */
int64_t gal_tree_words(gal_allocator *a, int tree, int partition, uint64_t *out,
                       uint64_t cap, uint32_t *num_layers, uint64_t *layer_words);

/* ---- global instance (global_allocator.cuh:41-97) ---- */
/* init_global_allocator / free_global_allocator / print_global_stats for the
 * library's own global_gallatin symbol (Gallatin<16 MiB, 16, 4096>). */
int gal_init_global(uint64_t num_bytes, uint64_t seed, int print_info);
int gal_free_global(void);
int gal_print_global_stats(void);
gal_allocator *gal_global_handle(void);

/* ---- workload kernels (the reference's test kernels) ---- */

/* One thread per request: out_dev[i] = allocator->malloc(sizes_dev[i]).
 * tests/src/gallatin_test.cu:59-86 alloc_one_size_pointer, variable sizes. */
int gal_malloc_bulk(gal_allocator *a, const uint64_t *sizes_dev, uint64_t n,
                    void **out_dev, uint32_t block_threads, void *stream);

/* One thread per pointer: allocator->free(ptrs_dev[i]) (NULL skipped).
 * tests/src/gallatin_test.cu:90-112 free_one_size_pointer. */
int gal_free_bulk(gal_allocator *a, void *const *ptrs_dev, uint64_t n,
                  uint32_t block_threads, void *stream);

/* Same through global_malloc / global_free (global_allocator.cuh:72-82). */
int gal_global_malloc_bulk(const uint64_t *sizes_dev, uint64_t n,
                           void **out_dev, uint32_t block_threads, void *stream);
int gal_global_free_bulk(void *const *ptrs_dev, uint64_t n,
                         uint32_t block_threads, void *stream);

/* Single-wave deterministic stream (one 64-lane wave).  Step layout: u32
 * kind (0 malloc, 1 free), u32 pad, u64 lane mask, u64 arg[64]; a malloc arg
 * is a size, a free arg is the index step*64+lane of an earlier malloc.
 * results_dev[step*64+lane] = byte offset from the pool base or ~0. */
typedef struct gal_wave_step {
  uint32_t kind;
  uint32_t pad;
  uint64_t mask;
  uint64_t arg[64];
} gal_wave_step;
int gal_run_wave_stream(gal_allocator *a, const gal_wave_step *steps_dev,
                        uint64_t nsteps, uint64_t *results_dev, void *stream);

/* Counters written by the workload kernels (device array, GAL_CTR_COUNT x
 * u64): [0] pairs (successful malloc+free), [1] misses (nullptr), [2]
 * pattern mismatches, [3] payload bytes written+read, [4] mallocs issued,
 * [5] bounded waits that ran out (batched streams), [6] frees of a table
 * entry that held nullptr (batched streams: the malloc had failed).
 * Batched streams also count [9] polls of frees whose entry was not yet
 * published.  The persistent stream kernel, when timed (GAL_BATCH_TIMED),
 * adds per wave wall-clock ticks (wall_clock64) spent [7] waiting for
 * entries, [10] in its malloc rounds and [11] freeing, and [8] its
 * lifetime, so [7] / [8] is the share of the kernel's wave time spent
 * waiting. */
enum { GAL_CTR_PAIRS = 0, GAL_CTR_MISSES = 1, GAL_CTR_BAD = 2,
       GAL_CTR_BYTES = 3, GAL_CTR_MALLOCS = 4, GAL_CTR_TIMEOUTS = 5,
       GAL_CTR_FREES_SKIPPED = 6, GAL_CTR_WAIT_TICKS = 7, GAL_CTR_WAVE_TICKS = 8,
       GAL_CTR_WAIT_POLLS = 9, GAL_CTR_MALLOC_TICKS = 10, GAL_CTR_FREE_TICKS = 11,
       GAL_CTR_COUNT = 12 };

/* Plain method: nthreads threads each malloc(size), write their uint32 id,
 * read it back and free (config "plain small"; gallatin_test.cu:59-112 in
 * one kernel). */
int gal_run_plain(gal_allocator *a, uint64_t nthreads, uint64_t size,
                  uint64_t *counters_dev, uint32_t block_threads, void *stream);

/* Mixed-size churn: nthreads threads x rounds of streams.h's churn with one
 * live allocation per thread, every allocation filled with its (thread,
 * round) pattern and verified before free (gallatin_churn.cu:37-95 with
 * log-uniform sizes). */
int gal_run_churn(gal_allocator *a, uint64_t seed, uint64_t nthreads,
                  uint32_t rounds, uint64_t *counters_dev,
                  uint32_t block_threads, void *stream);

/* Block and multi-segment batch, one step of streams.h's gal_big_* stream:
 * a launch in which n_small requests uniform in [64 KiB, 64 MiB] and n_large
 * uniform in [256 MiB, 2 GiB] each malloc and tag their first word, then a
 * launch that checks and frees them in gal_big_free_order.  ptrs_dev holds
 * n_small + n_large pointers (written by the first launch).  The request
 * shapes exercise malloc_block_allocation and malloc_segment_allocation
 * (gallatin.cuh:790-861) and the multi-segment free (gallatin.cuh:1378-1461). */
int gal_run_big(gal_allocator *a, uint64_t seed, uint32_t step, uint64_t n_small,
                uint64_t n_large, void **ptrs_dev, uint64_t *counters_dev, void *stream);

/* The same churn with one 64-bit word of each allocation written (its
 * (thread, round) pattern) and checked before free, as the reference's
 * churn test writes one word (gallatin_churn.cu:37-95): batched 0 runs the
 * per-thread path (`threads` per block), 1 the workgroup-batched path of
 * gal_run_churn_batched (`workgroups`, `threads` as there). */
int gal_run_churn_word(gal_allocator *a, uint64_t seed, uint64_t nthreads,
                       uint32_t rounds, uint64_t *counters_dev, uint32_t batched,
                       uint32_t workgroups, uint32_t threads, void *stream);

/* The same churn through the workgroup-batched path (batched.hpp): one
 * persistent launch of `workgroups` workgroups (0 = fill the device) of
 * `threads` threads (256, 512 or 1024; 0 = 1024) walks chunks of `threads`
 * stream threads; each round a workgroup frees its replaced allocations as
 * one batch and allocates as one batch.  No reference counterpart (the
 * reference serves warps, gallatin.cuh:865). */
int gal_run_churn_batched(gal_allocator *a, uint64_t seed, uint64_t nthreads,
                          uint32_t rounds, uint64_t *counters_dev,
                          uint32_t workgroups, uint32_t threads, void *stream);

/* ---- the host-pool global instance, combined, fused and poison ----
 * (global_allocator.cuh:44, 100-357.)  gal_init_global_host and friends
 * manage a second global instance, Gallatin<16 MiB, 16, 4096> with its pool
 * in pinned host memory (init_global_allocator_host, :100-104;
 * free_global_allocator_host, :106-112; print_global_stats_host, :123-131).
 * Bulk forms of the device calls, one thread per element:
 *  - fused (global_malloc_fused / global_free_fused, :195-216): the device
 *    pool first, the host pool when it has no room; free goes to the owner;
 *  - combined (global_malloc_combined / global_free_combined, :154-178):
 *    on_host picks the instance;
 *  - poison (global_malloc_poison / global_check_poison / global_free_poison,
 *    :238-357) on the device instance: size records before and after the
 *    allocation and a byte pattern in the slot's padding; check writes
 *    ok[i] = 1 when they are intact. */
int gal_init_global_host(uint64_t num_bytes, uint64_t seed, int print_info);
int gal_free_global_host(void);
int gal_print_global_stats_host(void);
gal_allocator *gal_global_host_handle(void);
int gal_global_malloc_fused_bulk(const uint64_t *sizes_dev, uint64_t n, void **out_dev,
                                 void *stream);
int gal_global_free_fused_bulk(void *const *ptrs_dev, uint64_t n, void *stream);
int gal_global_malloc_combined_bulk(const uint64_t *sizes_dev, uint64_t n, int on_host,
                                    void **out_dev, void *stream);
int gal_global_free_combined_bulk(void *const *ptrs_dev, uint64_t n, int on_host,
                                  void *stream);
int gal_global_malloc_poison_bulk(const uint64_t *sizes_dev, uint64_t n, void **out_dev,
                                  void *stream);
int gal_global_check_poison_bulk(void *const *ptrs_dev, uint64_t n, uint8_t *ok_dev,
                                 void *stream);
int gal_global_free_poison_bulk(void *const *ptrs_dev, uint64_t n, void *stream);

/* ---- allocator-backed data structures (on the global allocator) ----
 * gal_vector: fixed_vector<uint64_t, 64, 2^22> (data_structs/fixed_vector.cuh
 * :72-512, header include/gallatin/data_structs/fixed_vector.hpp).  insert
 * writes items[i] at a fresh position (out_positions[i], ~0 when full);
 * bulk != 0 takes the positions of a wave with one atomic (bulk_insert).
 * export copies items [0, size) to the host and returns size. */
typedef struct gal_vector gal_vector;
int gal_vector_create(gal_vector **out);
int gal_vector_destroy(gal_vector *v);
int gal_vector_insert(gal_vector *v, const uint64_t *items_dev, uint64_t n,
                      uint64_t *out_positions_dev, int bulk, void *stream);
int64_t gal_vector_export(gal_vector *v, uint64_t *out_host, uint64_t cap);

/* gal_table: extendible_hash_table<uint64_t, 0, ~0, uint64_t, 15, min, max>
 * (data_structs/extendible_ht.cuh:412-1437, header
 * include/gallatin/data_structs/extendible_ht.hpp); kind 0: 2^4 .. 2^20
 * buckets, kind 1: 2^16 .. 2^28.  Keys 0 and ~0 are reserved.  insert
 * replaces the value of a present key; remove has no reference
 * counterpart.  ok / found: one byte per key (may be NULL). */
typedef struct gal_table gal_table;
int gal_table_create(int kind, uint64_t seed, gal_table **out);
int gal_table_destroy(gal_table *t);
int gal_table_insert(gal_table *t, const uint64_t *keys_dev, const uint64_t *vals_dev,
                     uint64_t n, uint8_t *ok_dev, void *stream);
int gal_table_query(gal_table *t, const uint64_t *keys_dev, uint64_t n,
                    uint64_t *vals_dev, uint8_t *found_dev, void *stream);
int gal_table_remove(gal_table *t, const uint64_t *keys_dev, uint64_t n,
                     uint8_t *ok_dev, void *stream);
/* Keys present, buckets, deepest bucket depth, inserts refused.  Sync. */
int gal_table_stats(gal_table *t, uint64_t *keys, uint64_t *buckets,
                    uint64_t *level, uint64_t *failed);

/* gal_chain: chaining_table<Key, Val, size> (data_structs/chaining_table.cuh
 * :114-335, header include/gallatin/data_structs/chaining_table.hpp).
 * kind 0: 64-bit keys and values, 4 pairs per block (the reference's
 * default); 1: 64-bit, 16 pairs; 2: 32-bit, 4 pairs; 3: 32-bit, 32 pairs;
 * 4: 64-bit, 1 pair.  Keys and values are arrays of the kind's width.
 * create: nslots / size chains, key default_key marks an empty slot
 * (generate_on_device, chaining_table.cuh:132-156); prealloc != 0 gives
 * every chain its first block (generate_on_device_prealloc, :158-187).
 * insert (:240-282) stores a pair, duplicates included; ok[i] = 0 for
 * default_key or when no block could be allocated.  query (:284-309)
 * returns the first ready pair of the key in chain order.  stats: longest
 * chain and total blocks (print_chain_stats, :295-322), chains, refused
 * inserts; synchronous.  lengths: blocks of each chain, into nblocks
 * uint32 on the device.  destroy also returns the chain blocks. */
typedef struct gal_chain gal_chain;
int gal_chain_create(int kind, uint64_t nslots, uint64_t default_key, uint64_t seed,
                     int prealloc, gal_chain **out);
int gal_chain_destroy(gal_chain *c);
int gal_chain_insert(gal_chain *c, const void *keys_dev, const void *vals_dev, uint64_t n,
                     uint8_t *ok_dev, void *stream);
int gal_chain_query(gal_chain *c, const void *keys_dev, uint64_t n, void *vals_dev,
                    uint8_t *found_dev, void *stream);
int gal_chain_stats(gal_chain *c, uint64_t *max_len, uint64_t *total_blocks,
                    uint64_t *nblocks, uint64_t *failed);
int gal_chain_lengths(gal_chain *c, uint32_t *out_dev, void *stream);
int gal_chain_print(gal_chain *c);

/* gal_bqueue: block_queue<uint64_t, items_per_block> (data_structs/
 * block_queue.cuh:240-463, header include/gallatin/data_structs/
 * block_queue.hpp) on the global allocator; kind 0: 256 items per node (the
 * reference test's size), kind 1: 16.  enqueue (block_queue.cuh:331-405):
 * ok[i] = 0 when no node could be allocated.  dequeue has no working
 * reference counterpart (its code is commented out, :409-457): ok[i] = 0
 * when the queue was empty.  serial != 0: one thread does all n operations
 * in order.  reclaim returns the nodes that dequeues emptied (synchronous;
 * no kernel may use the queue).  stats: nodes in the list, retired nodes,
 * refused enqueues.  destroy returns every node. */
typedef struct gal_bqueue gal_bqueue;
int gal_bqueue_create(int kind, gal_bqueue **out);
int gal_bqueue_destroy(gal_bqueue *q);
int gal_bqueue_enqueue(gal_bqueue *q, const uint64_t *items_dev, uint64_t n,
                       uint8_t *ok_dev, int serial, void *stream);
int gal_bqueue_dequeue(gal_bqueue *q, uint64_t *out_dev, uint8_t *ok_dev, uint64_t n,
                       int serial, void *stream);
int gal_bqueue_reclaim(gal_bqueue *q, uint64_t *freed);
int gal_bqueue_stats(gal_bqueue *q, uint64_t *nodes, uint64_t *retired, uint64_t *failed);

/* gal_queue: queue<uint64_t, Gallatin<...>> (data_structs/queue.cuh:59-192,
 * header include/gallatin/data_structs/queue.hpp) on allocator `a`, or with
 * ring != 0 the ring of 8 such queues (circular_CAS_queue, queue.cuh:196).
 * enqueue: ok[i] = 0 when no node could be allocated.  dequeue: ok[i] = 0
 * when the (chosen) queue was empty.  serial != 0: one thread does all n
 * operations in order (FIFO checks).  destroy frees every node left. */
typedef struct gal_queue gal_queue;
int gal_queue_create(gal_allocator *a, int ring, gal_queue **out);
int gal_queue_destroy(gal_queue *q);
int gal_queue_enqueue(gal_queue *q, const uint64_t *items_dev, uint64_t n,
                      uint8_t *ok_dev, int serial, void *stream);
int gal_queue_dequeue(gal_queue *q, uint64_t *out_dev, uint8_t *ok_dev, uint64_t n,
                      int serial, void *stream);

/* ---- serving loop (no reference counterpart; on the global allocator) ----
 * Device-resident sessions, each owning a chunked_vector<uint64_t, 16,
 * 8192, 512> (include/gallatin/data_structs/chunked_vector.hpp: chunks of
 * 128 B doubling to 64 KiB, then up to 512 more 64 KiB chunks) and a shared
 * gal_table of kind 1.  A step is enqueued on `stream`: each live
 * session ends with probability end_q32 / 2^32 (vector freed) or appends
 * tokens x 32 words, tokens drawn from tokens_table[1024]; then `arrivals`
 * new sessions are admitted if the allocator has more than
 * reserve_segments free segments (else refused).  Alongside, on a second
 * stream of the serving object that starts after `stream`'s earlier work
 * and that `stream` waits for before the step ends, run ht_ops table
 * operations, 70 % insert / 30 % remove, keys uniform in [1, 2^28].
 * gate 0 counts free segments as the segment tree's set bits; gate 1 counts
 * the free memory in segments: those free segments plus the free blocks of
 * formatted segments (a block of class t is 2^t / blocks_per_segment of a
 * segment).
 * Counters (device, 16 x u64): [0] admitted, [1] refused, [2] ended, [3]
 * words appended, [4] words refused (a chunk could not be allocated, or
 * the vector is at its 4,210,688-word limit), [5] chunks allocated, [6]
 * chunk allocations failed, [7] chunks freed, [8] table inserts, [9]
 * removes that found the key, [10] removes of absent keys, [11] inserts
 * refused. */
typedef struct gal_serving gal_serving;
int gal_serving_create(uint32_t max_sessions, uint64_t seed,
                       const uint32_t *tokens_table_host, gal_serving **out);
int gal_serving_step(gal_serving *s, uint32_t step, uint32_t arrivals,
                     uint64_t reserve_segments, uint32_t gate, uint32_t end_q32,
                     uint32_t ht_ops, uint64_t *counters_dev, void *stream);
/* The gate's inputs at the last step: free segments, and free memory in
 * segments (gate 1's count).  Synchronous. */
int gal_serving_gate(gal_serving *s, uint64_t *free_segments, uint64_t *free_memory_segments);
/* Live sessions, words checked, words not equal to (id << 32) | position,
 * and the bytes of the chunks (and chunk indexes) the sessions hold.
 * Synchronous. */
int gal_serving_check(gal_serving *s, uint64_t *live, uint64_t *words, uint64_t *bad,
                      uint64_t *held_bytes);
/* Keys, buckets and refused inserts of the session table. */
int gal_serving_table(gal_serving *s, uint64_t *keys, uint64_t *buckets, uint64_t *failed);
/* Frees every live session's vector and the table. */
int gal_serving_destroy(gal_serving *s);

/* ---- batched request streams (no reference counterpart) ----
 * A stream is nbatches batches of ops; batch b is ops[offsets[b] ..
 * offsets[b+1]) (offsets is a host array of nbatches + 1 entries).  kind 0:
 * malloc(arg) into ptrs[slot]; kind 1: free(ptrs[slot]) (nullptr skipped),
 * where slot was filled by an earlier batch.  Each allocation holds its slot
 * number in its first 8 bytes, checked at free (counter [2]; an op with a
 * slot >= nslots or an unknown kind also counts there and is skipped).  One
 * persistent launch per batch, then one launch freeing whatever is left in
 * ptrs[0 .. nslots), so the table is empty again at the end.  batched: 0
 * plain per-thread malloc/free; 1 the workgroup-batched path (batched.hpp);
 * 2 mallocs through each workgroup's LDS slice cache (slice_cache.hpp) and
 * per-wave frees.  threads: 256, 512 or 1024 (0 = 1024; modes 1 and 2
 * only); workgroups: 0 = fill the device.  Counters as for the workload
 * kernels. */
typedef struct gal_batch_op {
  uint64_t arg;   /* malloc: bytes */
  uint32_t kind;  /* 0 malloc, 1 free */
  uint32_t slot;  /* pointer table entry */
} gal_batch_op;

int gal_run_batches(gal_allocator *a, const gal_batch_op *ops_dev,
                    const uint64_t *offsets, uint32_t nbatches, uint64_t *ptrs_dev,
                    uint64_t nslots, uint64_t *counters_dev, uint32_t threads,
                    uint32_t batched, uint32_t workgroups, void *stream);

/* Validity of a stream on the host (ops in host memory).  Every op has
 * kind 0 or 1 and a slot below nslots; a slot is filled again only after
 * its pointer was freed, and a free names a slot whose malloc ran in an
 * earlier batch.  persistent != 0 adds the requirement of the plan below:
 * each slot is filled at most once in the stream.  GAL_ERR_INVALID names
 * the first op that breaks a rule (gal_last_error).  No reference
 * counterpart. */
int gal_check_batch_stream(const gal_batch_op *ops_host, const uint64_t *offsets,
                           uint32_t nbatches, uint64_t nslots, int persistent);

/* The stream as one persistent launch, captured once into a hipGraph (on a
 * stream of its own: a reset kernel, the persistent kernel, the cleanup)
 * and replayed with gal_batch_plan_launch.  Workgroups take chunks of at
 * most 4 x threads ops (never across batches) in stream order from a
 * device queue and serve a chunk's mallocs before its frees; a free waits
 * only for its own table entry (bounded: a wait that runs out counts in
 * counter [5]).
 * REQUIREMENT: each slot is filled by at most one malloc in the whole
 * stream, and every free names a slot filled in an earlier batch.  The
 * chunks run at once, so a slot filled twice could be overwritten before
 * its first pointer is freed.  gal_batch_plan_create copies the ops to the
 * host once and refuses a stream that breaks this (gal_check_batch_stream
 * with persistent = 1) with GAL_ERR_INVALID.  The buffers must stay valid
 * until the plan is destroyed. */
/* Or-ed into `batched` of the persistent forms (gal_batch_plan_create,
 * gal_run_stream_static): the stream kernel reads the wall clock around its
 * phases and adds counters [7], [8], [10] and [11].  The reads cost about
 * 7 % of a replay, so untimed runs leave those counters at 0.  The eager
 * gal_run_batches ignores it. */
#define GAL_BATCH_TIMED 0x100u
typedef struct gal_batch_plan gal_batch_plan;
int gal_batch_plan_create(gal_allocator *a, const gal_batch_op *ops_dev,
                          const uint64_t *offsets, uint32_t nbatches,
                          uint64_t *ptrs_dev, uint64_t nslots,
                          uint64_t *counters_dev, uint32_t threads,
                          uint32_t batched, uint32_t workgroups,
                          gal_batch_plan **out);
int gal_batch_plan_launch(gal_batch_plan *plan, void *stream);

/* The persistent form of the stream run once, eagerly on `stream`, without
 * allocating anything: no counterpart in the reference.  It is for callers
 * that cannot allocate device memory (HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, requirements
 * (see gal_batch_plan_create; the slot rule is not checked here) and
 * results as gal_batch_plan_create + gal_batch_plan_launch.  The chunk
 * table, a copy of the batch offsets and the queue counter live in one
 * static device workspace of 2.5 MiB per device, held until the process
 * ends: at most 65,535 batches and 131,072 chunks of 4 x threads ops
 * (1,024 ops for batched = 0), else GAL_ERR_INVALID before any device work.
 * Every call on a device shares that workspace: calls that overlap on one
 * device race, so a caller orders them (one stream, or a synchronisation). */
int gal_run_stream_static(gal_allocator *a, const gal_batch_op *ops_dev,
                          const uint64_t *offsets, uint32_t nbatches, uint64_t *ptrs_dev,
                          uint64_t nslots, uint64_t *counters_dev, uint32_t threads,
                          uint32_t batched, uint32_t workgroups, void *stream);
int gal_batch_plan_destroy(gal_batch_plan *plan);

/* ---- measurement kernels (ceilings for the roofline) ---- */
/* n_ops relaxed agent-scope atomicAdd(u32) with the returned value kept:
 * op i targets word ((i / lanes_per_address) * stride) % span, so
 * lanes_per_address = 64 puts one address per wave and span = 1 one address
 * for the whole device. */
int gal_bench_atomic(uint32_t *words_dev, uint64_t span, uint64_t n_ops,
                     uint32_t lanes_per_address, uint32_t stride, void *stream);
/* Copy of `bytes` (multiple of 16) with 16-byte loads and stores, each
 * workgroup streaming contiguous 64 KiB tiles. */
int gal_bench_copy(const void *src_dev, void *dst_dev, uint64_t bytes,
                   void *stream);
/* Tree-search probe: about `waves` waves (rounded up to whole workgroups
 * of four) each run `searches` wave-cooperative successor searches from
 * pseudo-random starts, alternating between the instance's segment tree and
 * class tree `tree` of partition 0 (veb.cuh:878-916 search, read only). */
int gal_bench_tree(gal_allocator *a, uint32_t tree, uint32_t searches, uint32_t waves,
                   uint64_t *sink_dev, void *stream);
/* Copy of `bytes` cut into pieces of `piece_bytes` (both powers of two,
 * piece >= 16 B), each piece read from and written to pseudo-randomly
 * permuted positions: the bandwidth ceiling of payload traffic spread over
 * scattered allocations of that size. */
int gal_bench_scatter(const void *src_dev, void *dst_dev, uint64_t bytes,
                      uint64_t piece_bytes, void *stream);

/* Last error message of the calling thread (static storage). */
const char *gal_last_error(void);

#ifdef __cplusplus
}
#endif

#endif

/*
 * TEST INFRASTRUCTURE ONLY - the oracle (see gallatin_oracle.h).
 *
 * Sequential restatement of the Gallatin malloc/free policy.  Every function
 * cites the reference lines it follows (paths relative to the reference
 * checkout, include/gallatin/allocators/ unless stated otherwise).
 */
#define _GNU_SOURCE
#include "gallatin_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>

#include "../include/gallatin/streams.h"

#define BIT(i) (1ULL << (i))
#define LOWMASK(n) ((n) >= 64 ? ~0ULL : (BIT(n) - 1ULL))
#define SLICES_PER_BLOCK 4096u
#define TREE_SHIFT 20 /* block.cuh:53 GALLATIN_BLOCK_TREE_OFFSET */
#define COUNT_MASK ((1u << TREE_SHIFT) - 1u)
#define FREE_TREE 0xFFFFu
#define NULL_BLOCK 0xFFFFFFFFu

/* retry limits, gallatin.cuh:108-112 (defaults; orc_set_limits changes them) */
#define REQUEST_BLOCK_MAX_ATTEMPTS 1000
#define SLICE_ATTEMPTS (500 * 10) /* GALLATIN_MAX_ATTEMPTS x MALLOC_LOOP */
#define BLOCK_ATTEMPTS 500
#define SEGMENT_ATTEMPTS 500
#define SLOT_SCAN_CUTOFF 30 /* shared_block_storage.cuh:31 */
#define ORC_SPIN_CAP 1000000 /* oracle-only guard against unbounded loops */

/* ================================================================ utils */

/* alloc_utils.cuh:382-398 get_first_bit_bigger: ceil(log2(x)), x >= 1. */
int orc_clog2(uint64_t x) {
  return 63 - __builtin_clzll(x) + (__builtin_popcountll(x) != 1);
}

/* chaining_table.cuh:244 and :290: the chain of a key is MurmurHash64A over
 * the key's bytes (key_bytes = sizeof(Key), little-endian), modulo the
 * number of chains.  out[i] for keys[i], n keys packed back to back. */
void orc_chain_of(const void *keys, int key_bytes, uint64_t n, uint64_t seed,
                  uint64_t nchains, uint64_t *out) {
  const unsigned char *p = (const unsigned char *)keys;
  for (uint64_t i = 0; i < n; i++)
    out[i] = orc_murmur64a(p + i * (uint64_t)key_bytes, key_bytes, seed) % nchains;
}

/* The block and multi-segment request stream of streams.h (no reference
 * counterpart: the request shapes of the block + multi-segment workload),
 * evaluated by the same header the GPU kernels include: sizes[i] and the
 * request order[j] freed j-th in step `step`. */
void orc_big_stream(uint64_t seed, uint32_t step, uint64_t n_small, uint64_t n_large,
                    uint64_t *sizes, uint64_t *order) {
  const uint64_t n = n_small + n_large;
  for (uint64_t i = 0; i < n; i++) {
    if (sizes) sizes[i] = gal_big_size(seed, step, i, n_small);
    if (order) order[i] = gal_big_free_order(seed, step, i, n);
  }
}

/* alloc_utils.cuh:440-453 __cfcll: number of trailing one bits. */
int orc_cfcll(uint64_t bits) {
  int p = __builtin_popcountll(bits);
  if (p == 0 || p == 64) return p;
  return __builtin_ffsll((long long)~bits) - 1;
}

/* murmurhash.cuh:20-61, the published MurmurHash2 64-bit variant "64A". */
uint64_t orc_murmur64a(const void *key, int len, uint64_t seed) {
  const uint64_t m = 0xc6a4a7935bd1e995ULL;
  const int r = 47;
  uint64_t h = seed ^ ((uint64_t)len * m);
  const unsigned char *p = (const unsigned char *)key;
  int nblocks = len / 8;
  for (int i = 0; i < nblocks; i++) {
    uint64_t k;
    memcpy(&k, p + 8 * i, 8);
    k *= m;
    k ^= k >> r;
    k *= m;
    h ^= k;
    h *= m;
  }
  const unsigned char *tail = p + 8 * nblocks;
  switch (len & 7) {
    case 7: h ^= (uint64_t)tail[6] << 48; /* fall through */
    case 6: h ^= (uint64_t)tail[5] << 40; /* fall through */
    case 5: h ^= (uint64_t)tail[4] << 32; /* fall through */
    case 4: h ^= (uint64_t)tail[3] << 24; /* fall through */
    case 3: h ^= (uint64_t)tail[2] << 16; /* fall through */
    case 2: h ^= (uint64_t)tail[1] << 8;  /* fall through */
    case 1:
      h ^= (uint64_t)tail[0];
      h *= m;
  }
  h ^= h >> r;
  h *= m;
  h ^= h >> r;
  return h;
}

/* ============================================================= vEB tree */

/* Layer sizing, veb.cuh:595-618: num_layers = ceil(bitlen(U)/6); layer
 * L-1-i has universe u_i with u_{i+1} = ceil(u_i / 64). */
orc_veb *orc_veb_create(uint64_t universe, int filled) {
  if (universe == 0) return NULL;
  orc_veb *t = (orc_veb *)calloc(1, sizeof(orc_veb));
  int max_height = 64 - __builtin_clzll(universe);
  t->num_layers = (max_height - 1) / 6 + 1;
  t->universe = universe;
  uint64_t u = universe;
  for (int i = 0; i < t->num_layers; i++) {
    int l = t->num_layers - 1 - i;
    t->layer_universe[l] = u;
    t->layer_words[l] = (u - 1) / 64 + 1;
    t->bits[l] = (uint64_t *)calloc(t->layer_words[l], sizeof(uint64_t));
    /* init_bits, veb.cuh:72-84: every bit below the layer universe set. */
    if (filled)
      for (uint64_t b = 0; b < u; b++) t->bits[l][b / 64] |= BIT(b % 64);
    u = (u - 1) / 64 + 1;
  }
  return t;
}

void orc_veb_destroy(orc_veb *t) {
  if (!t) return;
  for (int l = 0; l < t->num_layers; l++) free(t->bits[l]);
  free(t);
}

int orc_veb_num_layers(const orc_veb *t) { return t->num_layers; }
uint64_t orc_veb_layer_words(const orc_veb *t, int l) { return t->layer_words[l]; }
uint64_t *orc_veb_layer_bits(orc_veb *t, int l) { return t->bits[l]; }

/* veb.cuh:737-755 remove: clear, float up while the word became empty. */
int orc_veb_remove(orc_veb *t, uint64_t x) {
  int layer = t->num_layers - 1;
  uint64_t high = x >> 6;
  int low = (int)(x & 63);
  uint64_t old = t->bits[layer][high];
  t->bits[layer][high] = old & ~BIT(low);
  if (!(old & BIT(low))) return 0;
  while (__builtin_popcountll(old) == 1) {
    /* float_up, veb.cuh:700-707 */
    layer -= 1;
    if (layer < 0) break;
    low = (int)(high & 63);
    high >>= 6;
    old = t->bits[layer][high];
    t->bits[layer][high] = old & ~BIT(low);
  }
  return 1;
}

/* veb.cuh:775-798 insert_force_update: set, float up while the word was
 * empty. */
int orc_veb_insert(orc_veb *t, uint64_t x) {
  int layer = t->num_layers - 1;
  uint64_t high = x >> 6;
  int low = (int)(x & 63);
  uint64_t old = t->bits[layer][high];
  t->bits[layer][high] = old | BIT(low);
  if (old & BIT(low)) return 0;
  while (__builtin_popcountll(old) == 0) {
    layer -= 1;
    if (layer < 0) break;
    low = (int)(high & 63);
    high >>= 6;
    old = t->bits[layer][high];
    t->bits[layer][high] = old | BIT(low);
  }
  return 1;
}

/* veb.cuh:824-829 query of the bottom layer. */
int orc_veb_query(const orc_veb *t, uint64_t x) {
  if (x >= t->universe) return 0;
  return (int)((t->bits[t->num_layers - 1][x >> 6] >> (x & 63)) & 1);
}

/* veb.cuh:250-275 layer::find_next: first set bit strictly above `low`
 * (low = -1: first set bit) in word `high`, -1 if none. */
static int layer_find_next(const orc_veb *t, int layer, uint64_t high, int low) {
  if (high >= t->layer_universe[layer]) return -1;
  if (high >= t->layer_words[layer]) return -1;
  uint64_t w = t->bits[layer][high];
  if (low == -1) return __builtin_ffsll((long long)w) - 1;
  return __builtin_ffsll((long long)(w & ~LOWMASK(low + 1))) - 1;
}

/* veb.cuh:878-916 successor_thorough: smallest set index strictly greater
 * than x, walking up to the first level with a later bit and back down. */
uint64_t orc_veb_successor(const orc_veb *t, uint64_t x) {
  uint64_t high = x >> 6;
  int low = (int)(x & 63);
  int layer = t->num_layers - 1;
  for (long spin = 0; spin < ORC_SPIN_CAP; spin++) {
    for (;;) {
      int found = layer_find_next(t, layer, high, low);
      if (found == -1) {
        if (layer == 0) return ORC_FAIL;
        layer -= 1;
        low = (int)(high & 63);
        high >>= 6;
        continue;
      }
      break;
    }
    while (layer != t->num_layers - 1) {
      low = layer_find_next(t, layer, high, low);
      if (low == -1) break;
      /* float_down, veb.cuh:709-715 */
      layer += 1;
      high = (high << 6) + (uint64_t)low;
      low = -1;
    }
    low = layer_find_next(t, layer, high, low);
    if (low != -1) return (high << 6) + (uint64_t)low;
  }
  return ORC_FAIL;
}

/* veb.cuh:940-944 find_first_valid_index. */
uint64_t orc_veb_find_first(const orc_veb *t) {
  if (orc_veb_query(t, 0)) return 0;
  return orc_veb_successor(t, 0);
}

/* veb.cuh:918-937 lock_offset(0) inside veb.cuh:979-991 malloc_first. */
uint64_t orc_veb_malloc_first(orc_veb *t) {
  for (int attempt = 0; attempt < 15; attempt++) {
    uint64_t start = 0;
    if (orc_veb_query(t, start) && orc_veb_remove(t, start)) return start;
    for (long spin = 0; spin < ORC_SPIN_CAP; spin++) {
      start = orc_veb_successor(t, start);
      if (start == ORC_FAIL) break;
      if (orc_veb_remove(t, start)) return start;
    }
  }
  return ORC_FAIL;
}

/* Parent-level maintenance after a contiguous change of the bottom layer,
 * veb.cuh:1100-1154 maybe_remove/return_layer_bits over the same word
 * ranges.  The parent bit of a child word is cleared when the word is empty
 * and set when it is non-empty (F1: the reference sets it only when the
 * child word is full, veb.cuh:1140). */
static void sync_range(orc_veb *t, uint64_t start_bit, uint64_t n) {
  for (int i = t->num_layers - 2; i >= 0; i--) {
    uint64_t hi = (start_bit + n - 1) / 64 + 1;
    uint64_t lo = start_bit / 64;
    start_bit = lo;
    n = hi - lo;
    for (uint64_t w = start_bit; w < start_bit + n; w++) {
      uint64_t child = t->bits[i + 1][w];
      if (child == 0)
        t->bits[i][w / 64] &= ~BIT(w % 64);
      else
        t->bits[i][w / 64] |= BIT(w % 64);
    }
  }
}

/* veb.cuh:323-452 layer::gather_multiple, sequential form: scan words from
 * the highest down for a word whose trailing-ones run ends a run of n set
 * bits that extends downward; take the lowest n bits of ... the top n bits
 * of that run when it fits in the word, else the trailing run plus full
 * words plus the top bits of the word below.  On a failed extension the
 * scan resumes at the word where it failed (veb.cuh:385-446).  F2: word 0
 * is a valid start and a run never extends below bit 0. */
static uint64_t layer_gather(orc_veb *t, uint64_t n) {
  int L = t->num_layers - 1;
  uint64_t *bits = t->bits[L];
  int64_t start_block = (int64_t)t->layer_words[L] - 1;
  while (start_block >= 0) {
    uint64_t word = bits[start_block];
    int contig = orc_cfcll(word);
    if (contig == 0) {
      start_block -= 1;
      continue;
    }
    if ((uint64_t)contig >= n) {
      uint64_t mask = LOWMASK(n) << (contig - (int)n);
      bits[start_block] &= ~mask;
      return (uint64_t)start_block * 64 + (uint64_t)(contig - (int)n);
    }
    /* contig < n: extend downward */
    uint64_t remaining = n - (uint64_t)contig;
    /* bits needed below word start_block */
    if ((uint64_t)start_block * 64 < remaining) return ORC_FAIL; /* F2 */
    int ok = 1;
    int64_t b = start_block;
    while (remaining > 64) {
      b--;
      if (bits[b] != ~0ULL) {
        ok = 0;
        break;
      }
      remaining -= 64;
    }
    if (!ok) {
      start_block = b; /* resume at the failing word */
      continue;
    }
    b--;
    uint64_t final_mask = LOWMASK(remaining) << (64 - remaining);
    if ((bits[b] & final_mask) != final_mask) {
      start_block = b;
      continue;
    }
    /* acquire: trailing run of the top word, full words, top bits below */
    bits[start_block] &= ~LOWMASK(contig);
    for (int64_t w = start_block - 1; w > b; w--) bits[w] = 0;
    bits[b] &= ~final_mask;
    return (uint64_t)b * 64 + 64 - remaining;
  }
  return ORC_FAIL;
}

/* veb.cuh:1160-1196 veb_tree::gather_multiple. */
uint64_t orc_veb_gather_multiple(orc_veb *t, uint64_t n) {
  if (n == 0) return ORC_FAIL;
  uint64_t found = layer_gather(t, n);
  if (found != ORC_FAIL) sync_range(t, found, n);
  return found;
}

/* veb.cuh:495-523 + 1198-1218 return_multiple (with F1). */
void orc_veb_return_multiple(orc_veb *t, uint64_t start, uint64_t n) {
  uint64_t *bits = t->bits[t->num_layers - 1];
  for (uint64_t i = start; i < start + n; i++) bits[i / 64] |= BIT(i % 64);
  sync_range(t, start, n);
}

uint64_t orc_veb_fill(const orc_veb *t) {
  uint64_t c = 0;
  int L = t->num_layers - 1;
  for (uint64_t w = 0; w < t->layer_words[L]; w++)
    c += (uint64_t)__builtin_popcountll(t->bits[L][w]);
  return c;
}

/* ============================================================ allocator */

typedef struct {
  uint32_t malloc_counter; /* (tree << 20) | claimed slices, block.cuh:59-68 */
  uint32_t free_counter;
} orc_block;

struct orc_alloc {
  uint64_t S, m, M;       /* bytes_per_segment, smallest, biggest */
  int T;                  /* num_trees, gallatin.cuh:367-370 */
  int smallest_bits;      /* clog2(m) */
  int block_tree_bits;    /* clog2(m * 4096) */
  uint64_t N;             /* num_segments, alloc_utils.cuh:223-226 */
  uint64_t B;             /* blocks_per_segment = S/(m*4096), memory_table.cuh:493 */
  orc_veb *segment_tree;
  orc_veb *sub_trees[ORC_MAX_TREES];
  uint16_t *chunk_ids;    /* memory_table.cuh:233 */
  int32_t *active;        /* memory_table.cuh:254 */
  uint32_t *q_malloc, *q_free, *q_final; /* memory_table.cuh:244-251 */
  uint32_t *ring;         /* queues, memory_table.cuh:241 (block ids) */
  orc_block *blocks;
  uint32_t slot_off[ORC_MAX_TREES + 1];
  uint32_t slot_count[ORC_MAX_TREES];
  uint32_t *slots;        /* pinned blocks, shared_block_storage.cuh:214-304 */
  uint64_t errors;
  unsigned char *host_pool; /* optional payload memory for churn replay */
  int lim_request, lim_slice, lim_block, lim_segment;
};

static uint64_t nb_of(const orc_alloc *a, int t) {
  /* memory_table.cuh:1057-1061 get_blocks_per_segment */
  return a->S / ((a->m << t) * SLICES_PER_BLOCK);
}

/* ---------------------------------------------------------- table ops */

/* memory_table.cuh:746-818 setup_segment. */
static int setup_segment(orc_alloc *a, uint64_t seg, int t) {
  if (a->chunk_ids[seg] != FREE_TREE) return 0; /* set_tree_id CAS, 823-828 */
  a->chunk_ids[seg] = (uint16_t)t;
  uint64_t nb = nb_of(a, t);
  for (uint64_t i = 0; i < nb; i++) a->ring[seg * a->B + i] = NULL_BLOCK;
  a->q_malloc[seg] = a->q_free[seg] = a->q_final[seg] = 0;
  a->active[seg] = (int32_t)nb - 1;
  return 1;
}

/* memory_table.cuh:915-993 get_block. */
static uint32_t get_block(orc_alloc *a, uint64_t seg, int t, int *empty) {
  *empty = 0;
  int32_t ac = a->active[seg]--;
  if (ac < 0) {
    a->active[seg]++;
    return NULL_BLOCK;
  }
  if (a->chunk_ids[seg] != (uint16_t)t) {
    a->active[seg]++;
    return NULL_BLOCK;
  }
  uint64_t nb = nb_of(a, t);
  uint32_t q = a->q_malloc[seg]++;
  uint32_t blk;
  if ((uint64_t)q < nb) {
    blk = (uint32_t)(seg * a->B + q);
  } else {
    uint64_t pos = seg * a->B + (q % nb);
    blk = a->ring[pos];
    a->ring[pos] = NULL_BLOCK;
  }
  if (blk == NULL_BLOCK) { /* memory_table.cuh:977-982 traps */
    a->errors++;
    return NULL_BLOCK;
  }
  a->blocks[blk].malloc_counter = (uint32_t)t << TREE_SHIFT; /* init_malloc */
  if (ac == 0) *empty = 1;
  return blk;
}

/* gallatin.cuh:1465-1510 gather_new_segment.  -1: no free segment,
 * -2: the segment could not be set up. */
static int64_t gather_new_segment(orc_alloc *a, int t) {
  uint64_t seg = orc_veb_malloc_first(a->segment_tree);
  if (seg == ORC_FAIL) return -1;
  if (!setup_segment(a, seg, t)) {
    a->errors++;
    return -2;
  }
  orc_veb_insert(a->sub_trees[t], seg);
  return (int64_t)seg;
}

/* gallatin.cuh:1525-1631 request_new_block_from_tree, deterministic segment
 * choice (find_first_valid_index).  The tree lock of 1513-1519 is always
 * free in a sequential run.  With no segment in the class tree and none free
 * the concurrent version polls for segments freed meanwhile; sequentially
 * every poll finds the same, so this returns at once. */
static uint32_t request_new_block(orc_alloc *a, int t) {
  int attempts = 0;
  long spins = 0;
  while (attempts < a->lim_request && spins < ORC_SPIN_CAP) {
    spins++;
    uint64_t seg = orc_veb_find_first(a->sub_trees[t]);
    if (seg == ORC_FAIL) {
      int64_t s = gather_new_segment(a, t);
      if (s == -1) return NULL_BLOCK;
      if (s < 0) {
        attempts++;
        continue;
      }
      seg = (uint64_t)s;
    }
    int last = 0;
    uint32_t blk = get_block(a, seg, t, &last);
    if (last) {
      orc_veb_remove(a->sub_trees[t], seg);
      gather_new_segment(a, t);
    }
    if (blk != NULL_BLOCK) return blk;
  }
  return NULL_BLOCK;
}

/* gallatin.cuh:1637-1726 return_block with memory_table.cuh:1158-1241
 * reserve_segment_slot / finish_freeing_block. */
static void return_block(orc_alloc *a, uint32_t blk, uint64_t seg, int t) {
  uint64_t nb = nb_of(a, t);
  uint32_t pos = a->q_free[seg]++;
  uint32_t r = (uint32_t)(pos % nb);
  a->ring[seg * a->B + r] = blk;
  a->q_final[seg] = pos + 1; /* finalize_free_queue */
  (void)r;
  /* REREGISTER_CUTOFF 0.1 (gallatin.cuh:1646) on the segment's fill (fix
   * F7): a return that brings the free blocks (active + 1) to at least
   * ceil(nb / 10), without completing the segment, inserts it into its class
   * tree when it is not there.  The reference compares the ring position
   * r = pos % nb with the cutoff instead (memory_table.cuh:1175-1182, "TODO
   * - make this actually calculate fill"). */
  if (!orc_veb_query(a->sub_trees[t], seg)) {
    int64_t f_after = (int64_t)a->active[seg] + 2;
    if (10 * f_after >= (int64_t)nb && f_after < (int64_t)nb)
      orc_veb_insert(a->sub_trees[t], seg);
  }
  int32_t rid = a->active[seg]++;
  int dereg = 0;
  if ((int64_t)rid == (int64_t)nb - 2) {
    if (a->active[seg] == (int32_t)nb - 1) {
      a->active[seg] = -1;
      dereg = 1;
    }
  }
  if (dereg) {
    orc_veb_remove(a->sub_trees[t], seg);
    if (a->chunk_ids[seg] == (uint16_t)t)
      a->chunk_ids[seg] = FREE_TREE;
    else
      a->errors++;
    orc_veb_insert(a->segment_tree, seg);
  }
}

/* gallatin.cuh:1732-1766 free_block. */
static void free_block(orc_alloc *a, uint32_t blk) {
  uint64_t seg = blk / a->B;
  int t = a->chunk_ids[seg];
  if (t >= a->T) {
    a->errors++;
    return;
  }
  return_block(a, blk, seg, t);
}

/* block.cuh:87-98 block_free_multiple + 200-212 reset_free, as used by
 * gallatin.cuh:1774-1827 free_offset. */
static void block_release(orc_alloc *a, uint32_t blk, uint32_t k) {
  uint32_t old = a->blocks[blk].free_counter;
  a->blocks[blk].free_counter = old + k;
  if (old + k == SLICES_PER_BLOCK) {
    a->blocks[blk].free_counter = 0;
    free_block(a, blk);
  }
}

/* ------------------------------------------------------- pinned slots */

/* shared_block_storage.cuh:119-134 get_valid_block_index, smid 0. */
static uint32_t valid_slot_index(const orc_alloc *a, int t) {
  uint32_t n = a->slot_count[t];
  uint32_t *s = a->slots + a->slot_off[t];
  int64_t original = 0;
  int64_t i = 0;
  int counter = 0;
  while (s[i] == NULL_BLOCK && i != original - 1) {
    i = (i + 1) % n;
    counter++;
    if (counter >= SLOT_SCAN_CUTOFF) break;
  }
  return (uint32_t)i;
}

/* gallatin.cuh:1083-1138 replace_block. */
static void replace_block(orc_alloc *a, int t, uint32_t slot, uint32_t blk) {
  uint32_t *s = a->slots + a->slot_off[t];
  if (s[slot] != blk) return; /* swap_out_block CAS lost */
  s[slot] = NULL_BLOCK;
  uint32_t nb = request_new_block(a, t);
  if (nb == NULL_BLOCK) return;
  if (s[slot] == NULL_BLOCK) {
    s[slot] = nb;
  } else {
    free_block(a, nb);
    a->errors++;
  }
}

/* --------------------------------------------------------------- boot */

void orc_destroy(orc_alloc *a) {
  if (!a) return;
  orc_veb_destroy(a->segment_tree);
  for (int t = 0; t < a->T; t++) orc_veb_destroy(a->sub_trees[t]);
  free(a->chunk_ids);
  free(a->active);
  free(a->q_malloc);
  free(a->q_free);
  free(a->q_final);
  free(a->ring);
  free(a->blocks);
  free(a->slots);
  if (a->host_pool) munmap(a->host_pool, a->N * a->S);
  free(a);
}

static int is_pow2(uint64_t x) { return x && !(x & (x - 1)); }

/* gallatin.cuh:324-421 generate_on_device + memory_table.cuh:471-647, with
 * the boot of boot_shared_block_container_one_thread (gallatin.cuh:195-221)
 * over slot counts from pinned_shared_blocks (shared_block_storage.cuh:251-274). */
orc_alloc *orc_create(uint64_t max_bytes, uint64_t S, uint64_t m, uint64_t M,
                      uint32_t slots0, uint32_t slots_min) {
  if (!is_pow2(S) || !is_pow2(m) || !is_pow2(M) || m > M) return NULL;
  if (S < M * SLICES_PER_BLOCK) return NULL; /* gallatin.cuh:300 */
  uint64_t N = max_bytes / S;
  if (N == 0) return NULL;
  orc_alloc *a = (orc_alloc *)calloc(1, sizeof(orc_alloc));
  a->S = S;
  a->m = m;
  a->M = M;
  a->T = orc_clog2(M) - orc_clog2(m) + 1;
  if (a->T > ORC_MAX_TREES) {
    free(a);
    return NULL;
  }
  a->smallest_bits = orc_clog2(m);
  a->block_tree_bits = orc_clog2(m * SLICES_PER_BLOCK);
  a->lim_request = REQUEST_BLOCK_MAX_ATTEMPTS;
  a->lim_slice = SLICE_ATTEMPTS;
  a->lim_block = BLOCK_ATTEMPTS;
  a->lim_segment = SEGMENT_ATTEMPTS;
  a->N = N;
  a->B = S / (m * SLICES_PER_BLOCK);
  a->segment_tree = orc_veb_create(N, 1);
  /* sub trees start full and boot_segment_trees (gallatin.cuh:136-145)
   * removes every index, leaving them empty */
  for (int t = 0; t < a->T; t++) {
    a->sub_trees[t] = orc_veb_create(N, 1);
    for (uint64_t i = 0; i < N; i++) orc_veb_remove(a->sub_trees[t], i);
  }
  a->chunk_ids = (uint16_t *)malloc(N * sizeof(uint16_t));
  a->active = (int32_t *)malloc(N * sizeof(int32_t));
  a->q_malloc = (uint32_t *)calloc(N, sizeof(uint32_t));
  a->q_free = (uint32_t *)calloc(N, sizeof(uint32_t));
  a->q_final = (uint32_t *)calloc(N, sizeof(uint32_t));
  a->ring = (uint32_t *)malloc(N * a->B * sizeof(uint32_t));
  a->blocks = (orc_block *)malloc(N * a->B * sizeof(orc_block));
  for (uint64_t i = 0; i < N; i++) {
    a->chunk_ids[i] = FREE_TREE;
    a->active[i] = -1; /* gallatin_init_counters_kernel, memory_table.cuh:173-201 */
  }
  for (uint64_t i = 0; i < N * a->B; i++) {
    a->ring[i] = NULL_BLOCK;
    a->blocks[i].malloc_counter = 4097; /* Block::init, block.cuh:64-68 */
    a->blocks[i].free_counter = 0;
  }
  uint32_t total = 0, p = slots0 ? slots0 : 1;
  for (int t = 0; t < a->T; t++) {
    a->slot_off[t] = total;
    a->slot_count[t] = p;
    total += p;
    p /= 2;
    if (p < slots_min) p = slots_min;
    if (p == 0) p = 1;
  }
  a->slot_off[a->T] = total;
  a->slots = (uint32_t *)malloc((size_t)total * sizeof(uint32_t));
  for (uint32_t i = 0; i < total; i++) a->slots[i] = NULL_BLOCK;
  for (int t = 0; t < a->T; t++) {
    for (uint32_t i = 0; i < a->slot_count[t]; i++) {
      uint32_t blk = request_new_block(a, t);
      if (blk == NULL_BLOCK) { /* boot_block prints and traps, 740-747 */
        orc_destroy(a);
        return NULL;
      }
      a->slots[a->slot_off[t] + i] = blk;
    }
  }
  return a;
}

void orc_set_limits(orc_alloc *a, int request, int slice, int block,
                    int segment) {
  if (request > 0) a->lim_request = request;
  if (slice > 0) a->lim_slice = slice;
  if (block > 0) a->lim_block = block;
  if (segment > 0) a->lim_segment = segment;
}

/* ------------------------------------------------------ address math */

/* memory_table.cuh:1072-1086 offset_to_allocation -> byte offset. */
static uint64_t offset_to_bytes(const orc_alloc *a, uint64_t off, int t) {
  uint64_t per_seg = a->S / a->m;
  uint64_t seg = off / per_seg;
  uint64_t rel = off % per_seg;
  return seg * a->S + rel * (a->m << t);
}

/* memory_table.cuh:1090-1127 allocation_to_offset. */
static uint64_t bytes_to_offset(const orc_alloc *a, uint64_t bytes, int t) {
  uint64_t seg = bytes / a->S;
  uint64_t in_seg = bytes - seg * a->S;
  return in_seg / (a->m << t) + seg * (a->S / a->m);
}

/* ------------------------------------------------------------ malloc */

enum { P_NONE = 0, P_SLICE, P_BLOCK, P_SEGMENT };

/* gallatin.cuh:1141-1147 get_tree_id_from_size and the size dispatch of
 * gallatin.cuh:1186-1276. */
static void classify(const orc_alloc *a, uint64_t size, int *path, int *cls,
                     uint32_t *count) {
  int t = (size < a->m) ? 0 : orc_clog2(size) - a->smallest_bits;
  *count = 1;
  if (t < a->T) {
    *path = P_SLICE;
    *cls = t;
    return;
  }
  int block_tree = orc_clog2(size) - a->block_tree_bits;
  if (block_tree < 0) {
    *path = P_SLICE;
    *count = 1u << (t - (a->T - 1));
    *cls = a->T - 1;
  } else if (block_tree < a->T) {
    *path = P_BLOCK;
    *cls = block_tree;
  } else {
    *path = P_SEGMENT;
    *cls = 0;
    *count = (uint32_t)((size - 1) / a->S + 1);
  }
}

/* Slice phase of one wave step: gallatin.cuh:865-1078 malloc_slice_allocation
 * with every coalesced group of the wave served per attempt. */
static void malloc_slices(orc_alloc *a, uint64_t mask, const int *cls,
                          const uint32_t *cnt, uint64_t *res) {
  int attempts[64] = {0};
  uint64_t pending = mask;
  while (pending) {
    uint32_t slot[64], blk[64], old[64], alloc_[64];
    int leader[64];
    uint64_t group[64];
    /* coalesced groups: labeled_partition by tree id (gallatin.cuh:883-885) */
    uint64_t rem = pending;
    while (rem) {
      int l = __builtin_ctzll(rem);
      uint64_t g = 0;
      for (int j = 0; j < 64; j++)
        if ((pending >> j & 1) && cls[j] == cls[l]) g |= BIT(j);
      for (int j = 0; j < 64; j++)
        if (g >> j & 1) {
          group[j] = g;
          leader[j] = l;
        }
      rem &= ~g;
    }
    /* leaders read their pinned slot (all before any claim) */
    for (int j = 0; j < 64; j++)
      if ((pending >> j & 1) && leader[j] == j) {
        slot[j] = valid_slot_index(a, cls[j]);
        blk[j] = a->slots[a->slot_off[cls[j]] + slot[j]];
      }
    /* F6: a group that finds no pinned block refills the start slot (0 in
     * this order) in lane order of the leaders; if no block can be had the
     * group's requests fail at once */
    for (int j = 0; j < 64; j++)
      if ((pending >> j & 1) && leader[j] == j && blk[j] == NULL_BLOCK) {
        uint32_t *sl = a->slots + a->slot_off[cls[j]];
        uint32_t cur = sl[0];
        if (cur == NULL_BLOCK) {
          uint32_t nb = request_new_block(a, cls[j]);
          if (nb != NULL_BLOCK) sl[0] = nb;
          cur = nb;
        }
        slot[j] = 0;
        blk[j] = cur;
        if (cur == NULL_BLOCK) pending &= ~group[j];
      }
    uint64_t have = 0;
    for (int j = 0; j < 64; j++)
      if (pending >> j & 1) {
        slot[j] = slot[leader[j]];
        blk[j] = blk[leader[j]];
        if (blk[j] != NULL_BLOCK) have |= BIT(j);
      }
    /* claims: exclusive scan + one atomicAdd per group (block.cuh:245-264) */
    uint32_t prefix[64];
    for (int j = 0; j < 64; j++)
      if ((have >> j & 1) && leader[j] == j) {
        uint32_t total = 0;
        for (int k = 0; k < 64; k++)
          if (group[j] >> k & 1) {
            prefix[k] = total;
            total += cnt[k];
          }
        old[j] = a->blocks[blk[j]].malloc_counter;
        a->blocks[blk[j]].malloc_counter = old[j] + total;
      }
    uint64_t replace_groups = 0, valid = 0;
    uint32_t leftover[64];
    for (int j = 0; j < 64; j++) {
      if (!(have >> j & 1)) continue;
      old[j] = old[leader[j]];
      uint32_t al = (old[j] & COUNT_MASK) + prefix[j];
      alloc_[j] = al;
      int start_valid = al <= 4095;
      int end_valid = al + cnt[j] <= 4096;
      if (start_valid && al + cnt[j] > 4095) replace_groups |= BIT(leader[j]);
      leftover[j] = (start_valid && end_valid) * (cnt[j] - 1) +
                    (start_valid && !end_valid) * (4096 - al);
      if (end_valid) valid |= BIT(j);
    }
    /* block_correct_frees (block.cuh:267-281), with F4 */
    uint64_t credit_free = 0;
    for (int j = 0; j < 64; j++)
      if ((have >> j & 1) && leader[j] == j) {
        uint32_t sum = 0;
        for (int k = 0; k < 64; k++)
          if (group[j] >> k & 1) sum += leftover[k];
        if (sum) {
          uint32_t of = a->blocks[blk[j]].free_counter;
          a->blocks[blk[j]].free_counter = of + sum;
          if (of + sum == SLICES_PER_BLOCK) credit_free |= BIT(j);
        }
      }
    for (int j = 0; j < 64; j++)
      if (credit_free >> j & 1) {
        a->blocks[blk[j]].free_counter = 0;
        free_block(a, blk[j]);
      }
    /* replace_block by the group leader (gallatin.cuh:1025-1033) */
    for (int j = 0; j < 64; j++)
      if (replace_groups >> j & 1) replace_block(a, cls[j], slot[j], blk[j]);
    /* check_valid (block.cuh:343-357) and results */
    for (int j = 0; j < 64; j++) {
      if (!(pending >> j & 1)) continue;
      if (valid >> j & 1) {
        uint64_t off = (uint64_t)alloc_[j] + (uint64_t)blk[j] * SLICES_PER_BLOCK;
        if ((old[j] >> TREE_SHIFT) == (uint32_t)cls[j]) {
          res[j] = offset_to_bytes(a, off, cls[j]);
          pending &= ~BIT(j);
          continue;
        }
        block_release(a, (uint32_t)(off / SLICES_PER_BLOCK), 1);
      }
      attempts[j]++;
      if (attempts[j] >= a->lim_slice) pending &= ~BIT(j);
    }
  }
}

/* gallatin.cuh:828-861 malloc_block_allocation + block.cuh:288-340
 * malloc_fill_block, retried per gallatin.cuh:1226-1231. */
static uint64_t malloc_block_path(orc_alloc *a, int bt) {
  for (int attempt = 0; attempt < a->lim_block; attempt++) {
    uint32_t blk = request_new_block(a, bt);
    if (blk == NULL_BLOCK) return ORC_FAIL; /* F6 */
    uint32_t merged = a->blocks[blk].malloc_counter;
    a->blocks[blk].malloc_counter = merged + 4096;
    uint32_t oldc = merged & COUNT_MASK;
    if ((merged >> TREE_SHIFT) != (uint32_t)bt) { /* check_valid -> trap */
      a->errors++;
      continue;
    }
    a->blocks[blk].free_counter += 4095 - oldc;
    if (oldc != 0) {
      block_release(a, blk, 1);
      continue;
    }
    return offset_to_bytes(a, (uint64_t)blk * SLICES_PER_BLOCK, bt);
  }
  return ORC_FAIL;
}

/* gallatin.cuh:790-826 malloc_segment_allocation, retried per 1246-1251. */
static uint64_t malloc_segment_path(orc_alloc *a, uint32_t n) {
  for (int attempt = 0; attempt < a->lim_segment; attempt++) {
    uint64_t idx = orc_veb_gather_multiple(a->segment_tree, n);
    if (idx == ORC_FAIL) return ORC_FAIL; /* F6 */
    uint32_t id = (uint32_t)a->T + 1 + n;
    if (a->chunk_ids[idx] != FREE_TREE || id >= FREE_TREE) {
      a->errors++;
      continue;
    }
    a->chunk_ids[idx] = (uint16_t)id;
    return idx * a->S;
  }
  return ORC_FAIL;
}

void orc_malloc_wave(orc_alloc *a, uint64_t mask, const uint64_t *sizes,
                     uint64_t *res) {
  int path[64], cls[64];
  uint32_t cnt[64];
  uint64_t slice_mask = 0;
  for (int j = 0; j < 64; j++) {
    res[j] = ORC_FAIL;
    if (!(mask >> j & 1)) continue;
    classify(a, sizes[j], &path[j], &cls[j], &cnt[j]);
    if (path[j] == P_SLICE) slice_mask |= BIT(j);
  }
  if (slice_mask) malloc_slices(a, slice_mask, cls, cnt, res);
  for (int j = 0; j < 64; j++)
    if ((mask >> j & 1) && path[j] == P_BLOCK) res[j] = malloc_block_path(a, cls[j]);
  for (int j = 0; j < 64; j++)
    if ((mask >> j & 1) && path[j] == P_SEGMENT)
      res[j] = malloc_segment_path(a, cnt[j]);
}

/* -------------------------------------------------------------- free */

/* gallatin.cuh:1378-1461 free + 1774-1827 free_offset. */
void orc_free_wave(orc_alloc *a, uint64_t mask, const uint64_t *offsets) {
  int tree[64];
  uint32_t blk[64];
  uint64_t slice_mask = 0;
  for (int j = 0; j < 64; j++) {
    if (!(mask >> j & 1)) continue;
    uint64_t off = offsets[j];
    if (off == ORC_FAIL || off >= a->N * a->S) { /* F5 */
      if (off != ORC_FAIL) a->errors++;
      tree[j] = -1;
      continue;
    }
    uint64_t seg = off / a->S;
    tree[j] = a->chunk_ids[seg];
    if (tree[j] == FREE_TREE || tree[j] == a->T) { /* F5 */
      a->errors++;
      tree[j] = -1;
      continue;
    }
    if (tree[j] < a->T) {
      blk[j] = (uint32_t)(bytes_to_offset(a, off, tree[j]) / SLICES_PER_BLOCK);
      slice_mask |= BIT(j);
    }
  }
  /* multi-segment frees in lane order: submit_segment_for_free
   * (gallatin.cuh:1283-1302) with F3 */
  for (int j = 0; j < 64; j++) {
    if (!(mask >> j & 1) || tree[j] <= a->T) continue;
    uint64_t seg = offsets[j] / a->S;
    uint64_t n = (uint64_t)(tree[j] - a->T - 1);
    a->chunk_ids[seg] = FREE_TREE;
    orc_veb_return_multiple(a->segment_tree, seg, n);
  }
  /* slice / block frees grouped by block id */
  uint64_t rem = slice_mask;
  while (rem) {
    int l = __builtin_ctzll(rem);
    uint64_t g = 0;
    for (int j = 0; j < 64; j++)
      if ((slice_mask >> j & 1) && blk[j] == blk[l]) g |= BIT(j);
    block_release(a, blk[l], (uint32_t)__builtin_popcountll(g));
    rem &= ~g;
  }
}

uint64_t orc_malloc(orc_alloc *a, uint64_t size) {
  uint64_t sizes[64] = {0}, res[64];
  sizes[0] = size;
  orc_malloc_wave(a, 1, sizes, res);
  return res[0];
}

void orc_free(orc_alloc *a, uint64_t offset) {
  uint64_t offs[64] = {0};
  offs[0] = offset;
  orc_free_wave(a, 1, offs);
}

int orc_run_stream(orc_alloc *a, const orc_wave_step *steps, uint64_t nsteps,
                   uint64_t *results) {
  for (uint64_t s = 0; s < nsteps; s++) {
    const orc_wave_step *st = &steps[s];
    uint64_t *out = results + s * 64;
    for (int j = 0; j < 64; j++) out[j] = ORC_FAIL;
    if (st->kind == 0) {
      orc_malloc_wave(a, st->mask, st->arg, out);
    } else if (st->kind == 1) {
      uint64_t offs[64];
      for (int j = 0; j < 64; j++) {
        offs[j] = ORC_FAIL;
        if (!(st->mask >> j & 1)) continue;
        uint64_t ref = st->arg[j];
        if (ref >= s * 64) return -1;
        offs[j] = results[ref];
      }
      orc_free_wave(a, st->mask, offs);
    } else {
      return -1;
    }
  }
  return 0;
}

/* ------------------------------------------------------------- drain */

/* Build-only quiescent drain: pinned blocks give back their unclaimed
 * slices (credit 4096 - claimed), then formatted segments whose blocks are
 * all available go back to the segment tree, as return_block would. */
void orc_drain(orc_alloc *a) {
  for (int t = 0; t < a->T; t++) {
    for (uint32_t i = 0; i < a->slot_count[t]; i++) {
      uint32_t *s = &a->slots[a->slot_off[t] + i];
      uint32_t blk = *s;
      if (blk == NULL_BLOCK) continue;
      *s = NULL_BLOCK;
      uint32_t c = a->blocks[blk].malloc_counter & COUNT_MASK;
      if (c > 4096) c = 4096;
      if (c < 4096) block_release(a, blk, 4096 - c);
    }
  }
  for (uint64_t seg = 0; seg < a->N; seg++) {
    int t = a->chunk_ids[seg];
    if (t >= a->T) continue;
    uint64_t nb = nb_of(a, t);
    if ((int64_t)a->active[seg] == (int64_t)nb - 1) {
      a->active[seg] = -1;
      orc_veb_remove(a->sub_trees[t], seg);
      a->chunk_ids[seg] = FREE_TREE;
      orc_veb_insert(a->segment_tree, seg);
    }
  }
}

/* -------------------------------------------------------- inspection */

void orc_get_geometry(const orc_alloc *a, orc_geometry *g) {
  g->num_segments = a->N;
  g->bytes_per_segment = a->S;
  g->smallest = a->m;
  g->biggest = a->M;
  g->blocks_per_segment = a->B;
  g->num_trees = (uint32_t)a->T;
  g->smallest_bits = (uint32_t)a->smallest_bits;
  g->total_slots = a->slot_off[a->T];
  g->pad = 0;
  g->errors = a->errors;
}

orc_veb *orc_segment_tree(orc_alloc *a) { return a->segment_tree; }
orc_veb *orc_sub_tree(orc_alloc *a, int t) {
  return (t >= 0 && t < a->T) ? a->sub_trees[t] : NULL;
}

void orc_segment_state(const orc_alloc *a, uint16_t *ids, int32_t *active) {
  for (uint64_t i = 0; i < a->N; i++) {
    if (ids) ids[i] = a->chunk_ids[i];
    if (active) active[i] = a->active[i];
  }
}

void orc_slots(const orc_alloc *a, uint32_t *out) {
  memcpy(out, a->slots, (size_t)a->slot_off[a->T] * sizeof(uint32_t));
}

void orc_blocks(const orc_alloc *a, uint32_t *out) {
  for (uint64_t i = 0; i < a->N * a->B; i++) {
    out[2 * i] = a->blocks[i].malloc_counter;
    out[2 * i + 1] = a->blocks[i].free_counter;
  }
}

/* ------------------------------------------------- sequential churn */

static unsigned char *pool_of(orc_alloc *a) {
  if (!a->host_pool) {
    void *p = mmap(NULL, a->N * a->S, PROT_READ | PROT_WRITE,
                   MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (p == MAP_FAILED) return NULL;
    a->host_pool = (unsigned char *)p;
  }
  return a->host_pool;
}

uint64_t orc_run_churn_sequential(orc_alloc *a, uint64_t seed, uint64_t tid0,
                                  uint64_t nthreads, uint32_t rounds,
                                  uint64_t *misses, uint64_t *bad) {
  unsigned char *pool = pool_of(a);
  uint64_t pairs = 0, miss = 0, wrong = 0;
  if (!pool) return 0;
  for (uint64_t tid = tid0; tid < tid0 + nthreads; tid++) {
    uint64_t live = ORC_FAIL, live_size = 0;
    uint32_t live_round = 0;
    for (uint32_t r = 0; r < rounds + 1; r++) {
      int replace = (r == 0) || (r == rounds) || gal_churn_coin(seed, tid, r);
      if (!replace) continue;
      if (live != ORC_FAIL) {
        uint64_t *w = (uint64_t *)(pool + live);
        uint64_t pat = gal_churn_pattern(tid, live_round);
        for (uint64_t i = 0; i < gal_fill_words(live_size); i++)
          if (w[i] != pat) wrong++;
        orc_free(a, live);
        pairs++;
        live = ORC_FAIL;
      }
      if (r == rounds) break;
      uint64_t size = gal_churn_size(seed, tid, r);
      live = orc_malloc(a, size);
      if (live == ORC_FAIL) {
        miss++;
        continue;
      }
      live_size = size;
      live_round = r;
      uint64_t *w = (uint64_t *)(pool + live);
      uint64_t pat = gal_churn_pattern(tid, r);
      for (uint64_t i = 0; i < gal_fill_words(size); i++) w[i] = pat;
    }
  }
  if (misses) *misses = miss;
  if (bad) *bad = wrong;
  return pairs;
}

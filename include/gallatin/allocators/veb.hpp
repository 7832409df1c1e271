// Layered 64-ary bitmap trees ("vEB" trees of the reference, veb.cuh:109-1228)
// laid out as 64-bit-word levels for wave64.
//
// Bit = 1 means available.  The bottom level holds one bit per segment; each
// upper bit summarises one 64-bit word of the level below.  Updates are
// single-lane atomics; searches are wave-cooperative: the k active lanes of
// the calling wave load k consecutive words of a level per pass and pick the
// first non-empty one with a ballot, so a 4096-segment level is one pass.
//
// Every search and update function that is not marked "per lane" must be
// called with the same arguments on every active lane (wave-uniform), and
// returns a wave-uniform result.
#ifndef GALLATIN_VEB_HPP
#define GALLATIN_VEB_HPP

#include "alloc_utils.hpp"
#include "layout.hpp"

namespace gallatin {
namespace allocators {
namespace veb {

using utils::atomic_and;
using utils::atomic_load;
using utils::atomic_or;

__device__ __forceinline__ uint64_t* word(const gal_tree& t, uint32_t layer,
                                          uint64_t w) {
  return t.words + t.layer_off[layer] + w;
}

// Re-establish "parent bit set <=> child word non-zero" above word `w` of
// `layer` after that word's emptiness changed (per lane).  Each writer of a
// parent bit re-reads the child afterwards and repeats on a mismatch, so
// once updates stop every parent bit equals the emptiness of its child word.
// Sequentially this is the float-up of veb.cuh:737-798 and the layer scans of
// veb.cuh:1100-1154 (with parent bits set for every non-empty child, where
// veb.cuh:1140 requires a full child).
__device__ inline void sync_up(const gal_tree& t, uint32_t layer, uint64_t w) {
  while (layer > 0) {
    uint64_t* child = word(t, layer, w);
    uint64_t* parent = word(t, layer - 1, w >> 6);
    const uint64_t pb = 1ULL << (w & 63);
    uint64_t old = 0;
    bool want = false;
    for (int guard = 0;; ++guard) {
      want = atomic_load(child) != 0;
      old = want ? atomic_or(parent, pb) : atomic_and(parent, ~pb);
      bool now = atomic_load(child) != 0;
      if (now == want || guard >= 64) break;
    }
    const bool had = (old & pb) != 0;
    if (had == want) return;
    const uint64_t nw = want ? (old | pb) : (old & ~pb);
    if ((old != 0) == (nw != 0)) return;
    --layer;
    w >>= 6;
  }
}

// insert_force_update (veb.cuh:775-798), per lane.  False if already set.
__device__ inline bool insert(const gal_tree& t, uint64_t x) {
  const uint32_t L = t.num_layers - 1;
  const uint64_t b = 1ULL << (x & 63);
  uint64_t old = atomic_or(word(t, L, x >> 6), b);
  if (old & b) return false;
  if (old == 0) sync_up(t, L, x >> 6);
  return true;
}

// remove (veb.cuh:737-755), per lane.  False if it was not set.
__device__ inline bool remove(const gal_tree& t, uint64_t x) {
  const uint32_t L = t.num_layers - 1;
  const uint64_t b = 1ULL << (x & 63);
  uint64_t old = atomic_and(word(t, L, x >> 6), ~b);
  if (!(old & b)) return false;
  if (old == b) sync_up(t, L, x >> 6);
  return true;
}

// query (veb.cuh:824-829), per lane.
__device__ inline bool query(const gal_tree& t, uint64_t x) {
  if (x >= t.universe) return false;
  return (atomic_load(word(t, t.num_layers - 1, x >> 6)) >> (x & 63)) & 1;
}

// Smallest available index >= x, or kFail.  Wave-cooperative: at each level
// the active lanes scan consecutive words; an empty scan climbs one level to
// skip to the next non-empty word, a hit descends.  Positions only increase,
// so the walk terminates even while other waves update the tree.  When the
// tree is quiescent this is find_first_valid_index (x = 0, veb.cuh:940-944)
// and query-or-successor_thorough (veb.cuh:878-916).
__device__ inline uint64_t find_ge(const gal_tree& t, uint64_t x) {
  const uint64_t act = wave::active();
  const uint32_t k = wave::count(act);
  const uint32_t rank = wave::count(act & wave::lanes_below());
  const uint32_t bottom = t.num_layers - 1;
  if (x >= t.universe) return kFail;
  uint32_t layer = bottom;
  uint64_t p = x;  // bit position within `layer`
  for (uint32_t guard = 0; guard < kSpinCap; ++guard) {
    const uint64_t nwords = t.layer_words[layer];
    const uint64_t w0 = p >> 6;
    if (w0 < nwords) {
      const uint64_t w = w0 + rank;
      uint64_t v = 0;
      if (w < nwords) {
        v = atomic_load(word(t, layer, w));
        if (w == w0) v &= ~0ULL << (p & 63);
      }
      const uint64_t nz = __ballot(v != 0);
      if (nz) {
        const int l = wave::first(nz);
        const uint64_t hv = wave::readlane(v, l);
        const uint64_t hw = wave::readlane(w, l);
        const uint64_t idx = (hw << 6) + (uint64_t)__builtin_ctzll(hv);
        if (layer == bottom) return idx;
        ++layer;
        p = idx << 6;
        continue;
      }
    }
    // nothing in [w0, w0 + k) of this level: continue after it one level up
    if (layer == 0) return kFail;
    --layer;
    p = w0 + k;
    if (p >= t.layer_words[layer] * 64) return kFail;
  }
  return kFail;
}

// True when the top level shows no available index (no atomics taken).
// While other waves update the tree the answer can be stale by one update,
// so callers use it only to skip work they would repeat anyway.
__device__ inline bool looks_empty(const gal_tree& t) {
  const uint64_t act = wave::active();
  const uint32_t k = wave::count(act);
  const uint32_t rank = wave::count(act & wave::lanes_below());
  const uint64_t nwords = t.layer_words[0];
  for (uint64_t w0 = 0; w0 < nwords; w0 += k) {
    const uint64_t w = w0 + rank;
    const uint64_t v = w < nwords ? atomic_load(word(t, 0, w)) : 0;
    if (__ballot(v != 0)) return false;
  }
  return true;
}

// malloc_first (veb.cuh:979-991): take the smallest available index.
__device__ inline uint64_t take_first(const gal_tree& t) {
  for (int attempt = 0; attempt < 15; ++attempt) {
    uint64_t x = find_ge(t, 0);
    while (x != kFail) {
      bool ok = false;
      if (wave::is_first()) ok = remove(t, x);
      if (wave::readfirst((uint32_t)ok)) return x;
      x = find_ge(t, x + 1);
    }
  }
  return kFail;
}

// Contiguous run of n available indices, claimed atomically; returns the
// lowest index or kFail.  Candidate runs end at the top of a word's
// trailing-ones run and extend downward (veb.cuh:323-452); the candidate of
// the highest word whose run is complete wins, which is the run the
// reference's top-down scan settles on when the tree is quiescent.  Word 0
// is a valid candidate and no run extends below index 0 (the reference
// skips word 0 and can index below it).  Each lane tests one candidate word
// per pass; the winning run is claimed word by word and rolled back if a
// concurrent update took part of it.
__device__ inline uint64_t gather_multiple(const gal_tree& t, uint64_t n) {
  const uint32_t L = t.num_layers - 1;
  const uint64_t nwords = t.layer_words[L];
  const uint64_t act = wave::active();
  const uint32_t k = wave::count(act);
  const uint32_t rank = wave::count(act & wave::lanes_below());
  if (n == 0 || n > t.universe) return kFail;
  for (int attempt = 0; attempt < 64; ++attempt) {
    uint64_t found = kFail;
    for (int64_t top = (int64_t)nwords - 1; top >= 0 && found == kFail;
         top -= (int64_t)k) {
      const int64_t w = top - (int64_t)rank;
      bool valid = false;
      uint32_t c = 0;
      if (w >= 0) {
        const uint64_t v = atomic_load(word(t, L, (uint64_t)w));
        c = (uint32_t)utils::cfcll(v);
        const uint64_t end = (uint64_t)w * 64 + c;
        if (c > 0 && end >= n) {
          valid = true;
          if (c < n) {
            uint64_t need = n - c;
            int64_t b = w;
            while (valid && need > 0) {
              --b;
              const uint64_t bv = atomic_load(word(t, L, (uint64_t)b));
              const uint64_t take = need >= 64 ? 64 : need;
              const uint64_t m = utils::low_mask(take) << (64 - take);
              valid = (bv & m) == m;
              need -= take;
            }
          }
        }
      }
      const uint64_t vm = __ballot(valid);
      if (vm) {
        const int l = wave::first(vm);  // lowest rank = highest word
        const uint64_t hw = wave::readlane((uint64_t)w, l);
        const uint64_t hc = wave::readlane(c, l);
        found = hw * 64 + hc - n;
      }
    }
    if (found == kFail) return kFail;
    // claim [found, found + n) word by word (first lane), roll back on loss
    uint32_t ok = 0;
    if (wave::is_first()) {
      ok = 1;
      uint64_t pos = found, left = n;
      uint64_t done_words = 0;
      while (left > 0) {
        const uint64_t wi = pos >> 6;
        const uint32_t bo = (uint32_t)(pos & 63);
        const uint64_t take = (64 - bo) < left ? (64 - bo) : left;
        const uint64_t m = utils::low_mask(take) << bo;
        const uint64_t old = atomic_and(word(t, L, wi), ~m);
        if ((old & m) != m) {
          // give back what we got here.  While our claim held the word at
          // zero, a concurrent remove of another bit may have cleared the
          // parent bit, so a give-back into an empty word re-syncs upward.
          if (old & m && atomic_or(word(t, L, wi), old & m) == 0) sync_up(t, L, wi);
          ok = 0;
          break;
        }
        ++done_words;
        pos += take;
        left -= take;
      }
      if (!ok) {
        // roll back the fully claimed words
        uint64_t rpos = found, rleft = n;
        for (uint64_t i = 0; i < done_words; ++i) {
          const uint64_t wi = rpos >> 6;
          const uint32_t bo = (uint32_t)(rpos & 63);
          const uint64_t take = (64 - bo) < rleft ? (64 - bo) : rleft;
          if (atomic_or(word(t, L, wi), utils::low_mask(take) << bo) == 0) sync_up(t, L, wi);
          rpos += take;
          rleft -= take;
        }
      } else {
        const uint64_t wlo = found >> 6, whi = (found + n - 1) >> 6;
        for (uint64_t wi = wlo; wi <= whi; ++wi)
          if (atomic_load(word(t, L, wi)) == 0) sync_up(t, L, wi);
      }
    }
    if (wave::readfirst(ok)) return found;
  }
  return kFail;
}

// return_multiple (veb.cuh:1198-1218), per lane.
__device__ inline void return_multiple(const gal_tree& t, uint64_t start,
                                       uint64_t n) {
  const uint32_t L = t.num_layers - 1;
  uint64_t pos = start, left = n;
  while (left > 0) {
    const uint64_t wi = pos >> 6;
    const uint32_t bo = (uint32_t)(pos & 63);
    const uint64_t take = (64 - bo) < left ? (64 - bo) : left;
    const uint64_t old = atomic_or(word(t, L, wi), utils::low_mask(take) << bo);
    if (old == 0) sync_up(t, L, wi);
    pos += take;
    left -= take;
  }
}

}  // namespace veb
}  // namespace allocators
}  // namespace gallatin

#endif

/*
 * Counter-based request-stream generators shared by the GPU workload kernels,
 * the host baselines and the tests.
 *
 * Every function is pure integer arithmetic keyed on (seed, thread id, round,
 * lane of use), so a request stream regenerated on the host is bit-identical
 * to the one a GPU kernel consumed.  Plain C99 so it can be included from C
 * (host baselines), C++ and HIP (where GAL_HD adds __host__ __device__).
 *
 * The request shapes follow the reference's own test kernels:
 *   - fixed-size allocate / write tid / read back / free
 *     (tests/src/gallatin_test.cu:59-112),
 *   - per-thread churn with a hash-chained size stream
 *     (tests/src/gallatin_churn.cu:37-95),
 * with the mixed-size churn drawing sizes log-uniformly instead of the
 * reference's `hash % max` (whose launch swaps min/max, gallatin_churn.cu:146,
 * so every request there is the maximum size).
 */
#ifndef GALLATIN_STREAMS_H
#define GALLATIN_STREAMS_H

#include <stdint.h>

#if defined(__HIPCC__)
#define GAL_HD __host__ __device__
#else
#define GAL_HD
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* splitmix64 finaliser: a bijective 64-bit mixer. */
static inline GAL_HD uint64_t gal_mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

/* Hash of (seed, a, b, c): independent streams per thread, round and use. */
static inline GAL_HD uint64_t gal_hash4(uint64_t seed, uint64_t a, uint64_t b,
                                        uint64_t c) {
  uint64_t h = gal_mix64(seed ^ 0x5851F42D4C957F2DULL);
  h = gal_mix64(h ^ a);
  h = gal_mix64(h ^ (b * 0xD6E8FEB86659FD93ULL));
  h = gal_mix64(h ^ (c * 0xA0761D6478BD642FULL));
  return h;
}

/* 2^f for a 32-bit fraction f (f / 2^32 in [0,1)), returned in Q30.
 * Horner evaluation of the Taylor series of e^(f ln 2) to 9 terms; integer
 * only, so host and device agree exactly.  Relative error < 2e-8. */
static inline GAL_HD uint64_t gal_exp2_frac_q30(uint32_t f) {
  const uint64_t ONE = 1ULL << 30;
  const uint64_t LN2_Q32 = 2977044472ULL; /* round(ln 2 * 2^32) */
  uint64_t y = ((uint64_t)f * LN2_Q32) >> 34; /* f*ln2 in Q30 */
  uint64_t t = ONE;
  for (int i = 9; i >= 1; --i) t = ONE + ((y * t) >> 30) / (uint64_t)i;
  return t;
}

/* Log-uniform size in [1, max_size): floor(max_size^u) with u uniform in
 * [0,1) taken from the top 32 bits of h; max_size must be a power of two
 * (4096 for the mixed churn). */
static inline GAL_HD uint64_t gal_log_uniform_size(uint64_t h, int log2_max) {
  uint64_t x = (h >> 32) * (uint64_t)log2_max; /* 32.32 fixed point, < log2_max */
  uint32_t k = (uint32_t)(x >> 32);
  uint32_t f = (uint32_t)x;
  uint64_t v = (gal_exp2_frac_q30(f) << k) >> 30;
  if (v < 1) v = 1;
  return v;
}

/* ---- mixed-size churn (log-uniform 1..4095 B, per-thread live slot) ---- */

#define GAL_CHURN_USE_SIZE 1u
#define GAL_CHURN_USE_COIN 2u

static inline GAL_HD uint64_t gal_churn_size(uint64_t seed, uint64_t tid,
                                             uint32_t round) {
  return gal_log_uniform_size(gal_hash4(seed, tid, round, GAL_CHURN_USE_SIZE),
                              12);
}

/* Round 0 always allocates; in rounds >= 1 a thread replaces its live
 * allocation when this coin is set (probability 1/2). */
static inline GAL_HD int gal_churn_coin(uint64_t seed, uint64_t tid,
                                        uint32_t round) {
  return (int)(gal_hash4(seed, tid, round, GAL_CHURN_USE_COIN) >> 63);
}

/* The uint64 pattern each allocation is filled with: (thread id, round). */
static inline GAL_HD uint64_t gal_churn_pattern(uint64_t tid, uint32_t round) {
  return (tid << 32) | (uint64_t)round;
}

/* Words filled/verified for a request of `size` bytes: ceil(size/8). */
static inline GAL_HD uint64_t gal_fill_words(uint64_t size) {
  return (size + 7) >> 3;
}

/* ---- block and multi-segment requests (one batch per step) ----
 * Requests 0 .. n_small-1 are uniform in [64 KiB, 64 MiB] (whole blocks of
 * the block classes, or 2-4 segment runs at 16 MiB segments); the rest are
 * uniform in [256 MiB, 2 GiB] (16-128 segment runs).  Every served request
 * gets a u64 tag in its first word, checked before it is freed, and the
 * frees of a step run in a pseudo-random order of the requests. */

#define GAL_BIG_USE_SIZE 3u
#define GAL_BIG_USE_ORDER 4u

static inline GAL_HD uint64_t gal_big_size(uint64_t seed, uint32_t step, uint64_t i,
                                           uint64_t n_small) {
  const uint64_t h = gal_hash4(seed, i, step, GAL_BIG_USE_SIZE);
  if (i < n_small) return (64ULL << 10) + h % ((64ULL << 20) - (64ULL << 10) + 1);
  return (256ULL << 20) + h % ((2ULL << 30) - (256ULL << 20) + 1);
}

static inline GAL_HD uint64_t gal_big_tag(uint32_t step, uint64_t i) {
  return gal_mix64(((uint64_t)step << 40) ^ i);
}

/* Request freed by the j-th free of a step: a bijection of [0, n), made by
 * walking the cycles of a bijection of [0, 2^k) (2^k >= n) until the value
 * falls below n.  The bijection is three rounds of xor with a constant,
 * multiplication by an odd constant mod 2^k and x ^= x >> ceil(k/2), each
 * invertible on k bits.  The walk ends: the cycle through j < n returns to
 * j itself at the latest. */
static inline GAL_HD uint64_t gal_big_free_order(uint64_t seed, uint32_t step, uint64_t j,
                                                 uint64_t n) {
  uint32_t k = 1;
  while ((1ULL << k) < n) ++k;
  const uint64_t mask = (1ULL << k) - 1;
  const uint64_t h = gal_hash4(seed, step, 0, GAL_BIG_USE_ORDER);
  const uint64_t g = gal_mix64(h);
  uint64_t x = j;
  do {
    for (uint32_t r = 0; r < 3; ++r) {
      const uint64_t c = (h >> (20 * r)) & mask;
      const uint64_t m = ((g >> (20 * r)) | 1) & mask;
      x = ((x ^ c) * m) & mask;
      x ^= x >> ((k + 1) / 2);
    }
  } while (x >= n);
  return x;
}

#ifdef __cplusplus
}
#endif

#endif /* GALLATIN_STREAMS_H */

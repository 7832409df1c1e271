// Host baselines: request streams of streams.h run through glibc malloc/free
// on the host cores, one std::thread per core.  The mixed-size churn gives
// each thread a contiguous shard of request-stream thread ids and does the
// same fill / verify / free work as the GPU churn kernel; the block and
// multi-segment batches do the same tag / check / free work as the GPU's.
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <unistd.h>

#include <atomic>
#include <cstdio>
#include <chrono>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gallatin/streams.h"

extern "C" {

typedef struct gal_cpu_result {
  uint64_t pairs, misses, bad, bytes;
  double seconds;
  uint32_t threads;
  uint32_t pad;
} gal_cpu_result;

int gal_cpu_online_cores(void) { return (int)sysconf(_SC_NPROCESSORS_ONLN); }

// Churn of stream threads [tid0, tid0 + n) for `rounds` rounds on `threads`
// host threads (0 = all online cores).
int gal_cpu_churn_glibc(uint64_t seed, uint64_t tid0, uint64_t n, uint32_t rounds,
                        uint32_t threads, gal_cpu_result* out) {
  if (!out) return -1;
  if (threads == 0) threads = (uint32_t)gal_cpu_online_cores();
  std::atomic<uint64_t> pairs{0}, misses{0}, bad{0}, bytes{0};
  auto work = [&](uint64_t lo, uint64_t hi) {
    uint64_t p = 0, m = 0, b = 0, by = 0;
    for (uint64_t tid = lo; tid < hi; ++tid) {
      uint64_t* live = nullptr;
      uint64_t live_words = 0;
      uint32_t live_round = 0;
      for (uint32_t r = 0; r <= rounds; ++r) {
        const bool replace = r == 0 || r == rounds || gal_churn_coin(seed, tid, r);
        if (!replace) continue;
        if (live) {
          const uint64_t pat = gal_churn_pattern(tid, live_round);
          for (uint64_t i = 0; i < live_words; ++i) b += live[i] != pat;
          free(live);
          ++p;
          by += live_words * 8;
          live = nullptr;
        }
        if (r == rounds) break;
        const uint64_t size = gal_churn_size(seed, tid, r);
        live = static_cast<uint64_t*>(malloc(size < 8 ? 8 : (size + 7) / 8 * 8));
        if (!live) {
          ++m;
          continue;
        }
        live_words = gal_fill_words(size);
        live_round = r;
        const uint64_t pat = gal_churn_pattern(tid, r);
        for (uint64_t i = 0; i < live_words; ++i) live[i] = pat;
        by += live_words * 8;
      }
    }
    pairs += p;
    misses += m;
    bad += b;
    bytes += by;
  };
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> pool;
  for (uint32_t i = 0; i < threads; ++i) {
    const uint64_t lo = tid0 + n * i / threads, hi = tid0 + n * (i + 1) / threads;
    pool.emplace_back(work, lo, hi);
  }
  for (auto& t : pool) t.join();
  const auto t1 = std::chrono::steady_clock::now();
  out->pairs = pairs;
  out->misses = misses;
  out->bad = bad;
  out->bytes = bytes;
  out->seconds = std::chrono::duration<double>(t1 - t0).count();
  out->threads = threads;
  out->pad = 0;
  return 0;
}

// Block and multi-segment batches of streams.h (gal_big_*), steps step0 ..
// step0 + steps - 1, on `threads` host threads that stay up for the whole
// run.  Each step: every thread mallocs a contiguous shard of the requests
// and tags each one's first word; after a barrier every thread checks and
// frees a contiguous shard of the free order; a second barrier ends the step.
int gal_cpu_big_glibc(uint64_t seed, uint32_t step0, uint32_t steps, uint64_t n_small,
                      uint64_t n_large, uint32_t threads, gal_cpu_result* out) {
  if (!out) return -1;
  if (threads == 0) threads = (uint32_t)gal_cpu_online_cores();
  const uint64_t n = n_small + n_large;
  std::vector<uint64_t*> ptrs(n, nullptr);
  std::atomic<uint64_t> pairs{0}, misses{0}, bad{0}, bytes{0};
  std::atomic<uint32_t> arrived{0}, phase{0};
  auto barrier = [&](uint32_t& local) {
    local ^= 1;
    if (arrived.fetch_add(1) + 1 == threads) {
      arrived.store(0);
      phase.store(local);
    } else {
      while (phase.load() != local) std::this_thread::yield();
    }
  };
  auto work = [&](uint32_t me) {
    uint64_t p = 0, m = 0, b = 0, by = 0;
    uint32_t local = 0;
    const uint64_t lo = n * me / threads, hi = n * (me + 1) / threads;
    for (uint32_t s = step0; s < step0 + steps; ++s) {
      for (uint64_t i = lo; i < hi; ++i) {
        uint64_t* q = static_cast<uint64_t*>(malloc(gal_big_size(seed, s, i, n_small)));
        if (q) {
          q[0] = gal_big_tag(s, i);
          by += 8;
        } else {
          ++m;
        }
        ptrs[i] = q;
      }
      barrier(local);
      for (uint64_t j = lo; j < hi; ++j) {
        const uint64_t i = gal_big_free_order(seed, s, j, n);
        uint64_t* q = ptrs[i];
        if (!q) continue;
        b += q[0] != gal_big_tag(s, i);
        by += 8;
        free(q);
        ++p;
      }
      barrier(local);
    }
    pairs += p;
    misses += m;
    bad += b;
    bytes += by;
  };
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> pool;
  for (uint32_t i = 0; i < threads; ++i) pool.emplace_back(work, i);
  for (auto& t : pool) t.join();
  const auto t1 = std::chrono::steady_clock::now();
  out->pairs = pairs;
  out->misses = misses;
  out->bad = bad;
  out->bytes = bytes;
  out->seconds = std::chrono::duration<double>(t1 - t0).count();
  out->threads = threads;
  out->pad = 0;
  return 0;
}

// Plain method: `launches` passes over thread ids [0, n), each doing
// malloc(16), writing its uint32 id, reading it back and freeing, on
// `threads` host threads (contiguous id shards).
int gal_cpu_plain_glibc(uint64_t n, uint32_t launches, uint32_t threads, gal_cpu_result* out) {
  if (!out) return -1;
  if (threads == 0) threads = (uint32_t)gal_cpu_online_cores();
  std::atomic<uint64_t> pairs{0}, misses{0}, bad{0};
  auto work = [&](uint64_t lo, uint64_t hi) {
    uint64_t p = 0, m = 0, b = 0;
    for (uint32_t l = 0; l < launches; ++l) {
      for (uint64_t tid = lo; tid < hi; ++tid) {
        volatile uint32_t* q = static_cast<volatile uint32_t*>(malloc(16));
        if (!q) {
          ++m;
          continue;
        }
        q[0] = (uint32_t)tid;
        b += q[0] != (uint32_t)tid;
        free((void*)q);
        ++p;
      }
    }
    pairs += p;
    misses += m;
    bad += b;
  };
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> pool;
  for (uint32_t i = 0; i < threads; ++i)
    pool.emplace_back(work, n * i / threads, n * (i + 1) / threads);
  for (auto& t : pool) t.join();
  const auto t1 = std::chrono::steady_clock::now();
  out->pairs = pairs;
  out->misses = misses;
  out->bad = bad;
  out->bytes = 8 * pairs;
  out->seconds = std::chrono::duration<double>(t1 - t0).count();
  out->threads = threads;
  out->pad = 0;
  return 0;
}

// A batched request stream (the layout of gal_batch_op in gallatin_capi.h):
// batch b is ops[offsets[b] .. offsets[b + 1]); kind 0 mallocs arg bytes
// into table entry `slot` and tags its first word with the slot number,
// kind 1 checks and frees the entry's pointer.  The ops of a batch are
// split over `threads` host threads in contiguous shards, with a barrier
// between batches (a free only names an entry of an earlier batch); what
// is left in the table is freed at the end.
typedef struct gal_cpu_batch_op {
  uint64_t arg;
  uint32_t kind;
  uint32_t slot;
} gal_cpu_batch_op;

int gal_cpu_batches_glibc(const gal_cpu_batch_op* ops, const uint64_t* offsets,
                          uint32_t nbatches, uint64_t nslots, uint32_t threads,
                          gal_cpu_result* out) {
  if (!out || !offsets || (nbatches && !ops)) return -1;
  if (threads == 0) threads = (uint32_t)gal_cpu_online_cores();
  std::vector<uint64_t*> table(nslots, nullptr);
  std::atomic<uint64_t> pairs{0}, misses{0}, bad{0};
  std::atomic<uint32_t> arrived{0}, phase{0};
  auto barrier = [&](uint32_t& local) {
    local ^= 1;
    if (arrived.fetch_add(1) + 1 == threads) {
      arrived.store(0);
      phase.store(local);
    } else {
      while (phase.load() != local) std::this_thread::yield();
    }
  };
  auto work = [&](uint32_t me) {
    uint64_t p = 0, m = 0, b = 0;
    uint32_t local = 0;
    for (uint32_t bt = 0; bt < nbatches; ++bt) {
      const uint64_t n = offsets[bt + 1] - offsets[bt];
      const uint64_t lo = offsets[bt] + n * me / threads;
      const uint64_t hi = offsets[bt] + n * (me + 1) / threads;
      for (uint64_t i = lo; i < hi; ++i) {
        const gal_cpu_batch_op& op = ops[i];
        if (op.slot >= nslots) {
          ++b;
          continue;
        }
        if (op.kind == 0) {
          uint64_t* q = static_cast<uint64_t*>(malloc(op.arg < 8 ? 8 : op.arg));
          if (!q) {
            ++m;
            continue;
          }
          q[0] = op.slot;
          table[op.slot] = q;
        } else {
          uint64_t* q = table[op.slot];
          table[op.slot] = nullptr;
          if (!q) continue;
          b += q[0] != op.slot;
          free(q);
          ++p;
        }
      }
      barrier(local);
    }
    for (uint64_t i = nslots * me / threads; i < nslots * (me + 1) / threads; ++i) {
      if (!table[i]) continue;
      b += table[i][0] != i;
      free(table[i]);
      table[i] = nullptr;
      ++p;
    }
    pairs += p;
    misses += m;
    bad += b;
  };
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> pool;
  for (uint32_t i = 0; i < threads; ++i) pool.emplace_back(work, i);
  for (auto& t : pool) t.join();
  const auto t1 = std::chrono::steady_clock::now();
  out->pairs = pairs;
  out->misses = misses;
  out->bad = bad;
  out->bytes = 16 * pairs;
  out->seconds = std::chrono::duration<double>(t1 - t0).count();
  out->threads = threads;
  out->pad = 0;
  return 0;
}

// The serving loop on host threads: the step of the device loop with glibc
// malloc/free for the session vectors and a std::unordered_map sharded by
// key hash (one shard per thread, so no locks) for the table.
//   1. each live session ends with probability p_end (its chunks are freed)
//      or appends tokens x 32 words (tokens from the 1024-entry quantile
//      table) to a vector of chunks of 16, 16, 32, ... 8192 words, then
//      further chunks of 8192 words (the device's chunked_vector<uint64_t,
//      16, 8192, 512>); word p of session id holds (id << 32) | p;
//   2. Poisson(arrivals_mean) arrivals take free slots while the chunk
//      bytes held stay below (1 - reserve_fraction) x pool_bytes (the
//      device gate counts free segments of the same pool instead);
//   3. ht_ops table operations, 70 % insert / 30 % remove of keys uniform
//      in [1, 2^28].
// Steps 0 .. warm-1 are untimed; step_ms receives the durations of the
// `steps` timed steps.
typedef struct gal_cpu_serving_result {
  uint64_t appended_words, ended, table_ops, table_keys, bad;
  uint64_t admitted, refused, held_bytes;
  uint32_t live, threads;
} gal_cpu_serving_result;

namespace {
inline uint64_t srv_mix(uint64_t a, uint64_t b, uint64_t c) {
  uint64_t x = a * 0x9E3779B97F4A7C15ULL ^ b;
  x ^= x >> 31;
  x *= 0xBF58476D1CE4E5B9ULL;
  x ^= c + 0x94D049BB133111EBULL;
  x ^= x >> 29;
  x *= 0x94D049BB133111EBULL;
  return x ^ (x >> 32);
}
constexpr uint32_t kSrvGrow = 11;          // chunks 0..10: 16, 16, 32, ... 8192 words
constexpr uint64_t kSrvGrowWords = 16384;  // words in the doubling chunks
constexpr uint64_t kSrvTail = 8192;        // words per later chunk
constexpr uint64_t kSrvMaxWords = kSrvGrowWords + 512 * kSrvTail;
inline uint32_t srv_chunk_of(uint64_t p) {
  if (p >= kSrvGrowWords) return kSrvGrow + (uint32_t)((p - kSrvGrowWords) / kSrvTail);
  return p < 16 ? 0u : (uint32_t)(63 - __builtin_clzll(p >> 4)) + 1u;
}
inline uint64_t srv_chunk_start(uint32_t d) {
  return d == 0 ? 0 : (d < kSrvGrow ? 16ULL << (d - 1) : kSrvGrowWords + (d - kSrvGrow) * kSrvTail);
}
inline uint64_t srv_chunk_words(uint32_t d) {
  return d <= 1 ? 16 : (d < kSrvGrow ? 16ULL << (d - 1) : kSrvTail);
}
struct cpu_session {
  std::vector<uint64_t*> chunk;
  uint64_t id;
  uint64_t size;
  uint32_t alive;
};
}  // namespace

int gal_cpu_serving_glibc(uint64_t seed, uint32_t warm, uint32_t steps, uint32_t max_sessions,
                          double arrivals_mean, double p_end, uint32_t ht_ops,
                          const uint32_t* tokens, uint64_t pool_bytes, double reserve_fraction,
                          uint32_t threads, double* step_ms, gal_cpu_serving_result* out) {
  if (!out || !tokens || !step_ms || max_sessions == 0) return -1;
  if (threads == 0) threads = (uint32_t)gal_cpu_online_cores();
  std::vector<cpu_session> ss(max_sessions);
  for (auto& s : ss) s = cpu_session{{}, 0, 0, 0};
  std::vector<uint32_t> free_slots(max_sessions);
  for (uint32_t i = 0; i < max_sessions; ++i) free_slots[i] = max_sessions - 1 - i;
  std::vector<std::unordered_map<uint64_t, uint64_t>> shard(threads);
  const uint32_t end_q32 = (uint32_t)(p_end * 4294967296.0);
  const double limit = (1.0 - reserve_fraction) * (double)pool_bytes;
  std::mt19937_64 rng(seed);
  std::poisson_distribution<uint32_t> arrivals(arrivals_mean);
  uint64_t next_id = 0, admitted = 0, refused = 0;
  std::atomic<uint64_t> words{0}, ended{0}, bad{0};
  std::atomic<int64_t> held{0};
  std::vector<std::vector<uint32_t>> ended_slots(threads);
  for (uint32_t step = 0; step < warm + steps; ++step) {
    const auto t0 = std::chrono::steady_clock::now();
    // 1. session pass, contiguous slot shards
    std::vector<std::thread> pool;
    for (uint32_t me = 0; me < threads; ++me) {
      pool.emplace_back([&, me]() {
        uint64_t w = 0, e = 0, b = 0;
        int64_t h = 0;
        ended_slots[me].clear();
        const uint32_t lo = (uint32_t)((uint64_t)max_sessions * me / threads);
        const uint32_t hi = (uint32_t)((uint64_t)max_sessions * (me + 1) / threads);
        for (uint32_t i = lo; i < hi; ++i) {
          cpu_session& s = ss[i];
          if (!s.alive) continue;
          const uint64_t r = srv_mix(seed, s.id, step);
          if ((uint32_t)r < end_q32) {
            if (!s.chunk.empty()) b += s.chunk[0][0] != (s.id << 32);  // its first word
            for (uint32_t d = 0; d < s.chunk.size(); ++d) {
              free(s.chunk[d]);
              h -= (int64_t)(srv_chunk_words(d) * 8);
            }
            s.chunk.clear();
            s.alive = 0;
            ended_slots[me].push_back(i);
            ++e;
            continue;
          }
          const uint64_t want = (uint64_t)tokens[(r >> 32) & 1023] * 32;
          uint64_t at = s.size;
          const uint64_t end = at + want > kSrvMaxWords ? kSrvMaxWords : at + want;
          while (at < end) {
            const uint32_t d = srv_chunk_of(at);
            const uint64_t lo_d = srv_chunk_start(d), n_d = srv_chunk_words(d);
            if (d >= s.chunk.size()) {
              s.chunk.push_back(static_cast<uint64_t*>(malloc(n_d * 8)));
              h += (int64_t)(n_d * 8);
            }
            uint64_t* c = s.chunk[d];
            const uint64_t stop = lo_d + n_d < end ? lo_d + n_d : end;
            for (uint64_t p = at; p < stop; ++p) c[p - lo_d] = (s.id << 32) | p;
            w += stop - at;
            at = stop;
          }
          s.size = at;
        }
        words += w;
        ended += e;
        bad += b;
        held += h;
      });
    }
    for (auto& t : pool) t.join();
    pool.clear();
    for (auto& v : ended_slots)
      for (uint32_t i : v) free_slots.push_back(i);
    // 2. admission, gated on the bytes held
    const uint32_t a = arrivals(rng);
    const bool room = (double)held.load() < limit;
    for (uint32_t k = 0; k < a; ++k) {
      if (!room || free_slots.empty()) {
        if (step >= warm) ++refused;
        continue;
      }
      const uint32_t i = free_slots.back();
      free_slots.pop_back();
      ss[i].chunk.clear();
      ss[i].id = next_id++;
      ss[i].size = 0;
      ss[i].alive = 1;
      if (step >= warm) ++admitted;
    }
    // 3. table traffic, each thread serving the keys of its shard
    for (uint32_t me = 0; me < threads; ++me) {
      pool.emplace_back([&, me]() {
        auto& m = shard[me];
        for (uint32_t i = 0; i < ht_ops; ++i) {
          const uint64_t hh = srv_mix(seed ^ 0x7A11E5ULL, step, i);
          const uint64_t key = (hh & ((1ULL << 28) - 1)) + 1;
          if (srv_mix(key, 0, 0) % threads != me) continue;
          if ((hh >> 32) % 10 < 7)
            m[key] = hh;
          else
            m.erase(key);
        }
      });
    }
    for (auto& t : pool) t.join();
    const auto t1 = std::chrono::steady_clock::now();
    const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    if (step >= warm) step_ms[step - warm] = ms;
    if ((step + 1) % 50 == 0 || step + 1 == warm + steps) {
      fprintf(stderr, "cpu serving: step %u of %u, %.1f ms, %.1f GB held, %zu free slots\n",
              step + 1, warm + steps, ms, (double)held.load() / 1e9, free_slots.size());
      fflush(stderr);
    }
  }
  uint32_t live = 0;
  for (auto& s : ss) {
    live += s.alive;
    for (auto* c : s.chunk) free(c);
  }
  uint64_t keys = 0;
  for (auto& m : shard) keys += m.size();
  out->appended_words = words;
  out->ended = ended;
  out->table_ops = (uint64_t)ht_ops * (warm + steps);
  out->table_keys = keys;
  out->bad = bad;
  out->admitted = admitted;
  out->refused = refused;
  out->held_bytes = (uint64_t)held.load();
  out->live = live;
  out->threads = threads;
  return 0;
}

}  // extern "C"

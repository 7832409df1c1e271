// Host baseline: the mixed-size churn request stream of streams.h run through
// glibc malloc/free on the host cores, one std::thread per core, each thread
// taking a contiguous shard of request-stream thread ids and doing the same
// fill / verify / free work as the GPU churn kernel.
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <unistd.h>

#include <atomic>
#include <chrono>
#include <thread>
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

}  // extern "C"

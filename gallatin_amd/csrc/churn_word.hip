// The mixed-size churn with one 64-bit word of each allocation written and
// checked, the pattern of the reference's churn test
// (tests/src/gallatin_churn.cu:37-95 writes the thread id into the first
// word with atomicExch and checks it before the free).  Same request stream
// as k_churn / k_churn_batched (streams.h): sizes, coins and patterns; only
// the payload differs, so the pair rate is set by the allocator rather than
// by HBM.  Kept in a file of its own so that the churn kernels of the
// stated workload compile exactly as they are measured.
#include <hip/hip_runtime.h>

#include "gallatin/allocators/batched.hpp"
#include "gallatin/allocators/gallatin.hpp"
#include "gallatin/streams.h"
#include "gallatin_capi.h"
#include "internal.hpp"

using namespace gallatin;
using namespace gallatin::allocators;
using namespace gallatin_capi;

namespace {

#include "lane_counts.hpp"

// A replaced allocation: its word checked, then its pair counted.
__device__ __forceinline__ void check_word(const uint64_t* live, uint64_t pat, lane_counts& c) {
  c.bad += *live != pat;
  c.pairs += 1;
  c.bytes += 8;
}

// A new allocation: its word written.
__device__ __forceinline__ void write_word(uint64_t* np, uint64_t pat, lane_counts& c) {
  *np = pat;
  c.bytes += 8;
}

__global__ void k_churn_word(gal_state* s, uint64_t seed, uint64_t n, uint32_t rounds,
                             uint64_t* ctr) {
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = tid < n;
  uint64_t* live = nullptr;
  uint32_t live_round = 0;
  lane_counts c;
  for (uint32_t r = 0; r <= rounds; ++r) {
    const bool replace =
        active && (r == 0 || r == rounds || gal_churn_coin(seed, tid, r) != 0);
    if (replace && live != nullptr) {
      check_word(live, gal_churn_pattern(tid, live_round), c);
      detail::free_impl(s, live);
      live = nullptr;
    }
    if (r == rounds) break;
    if (replace) {
      uint64_t* np =
          static_cast<uint64_t*>(detail::malloc_impl(s, gal_churn_size(seed, tid, r)));
      c.mallocs += 1;
      if (np) {
        write_word(np, gal_churn_pattern(tid, r), c);
        live = np;
        live_round = r;
      } else {
        c.misses += 1;
      }
    }
  }
  add_counters(ctr, c);
}

template <int kThreads>
__global__ __launch_bounds__(kThreads) void k_churn_batched_word(gal_state* s, uint64_t seed,
                                                                 uint64_t n, uint32_t rounds,
                                                                 uint64_t* ctr) {
  __shared__ batch::lds_state lds;
  batch::batch_init(lds);
  lane_counts c;
  const uint64_t nchunks = (n + blockDim.x - 1) / blockDim.x;
  for (uint64_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint64_t tid = chunk * blockDim.x + threadIdx.x;
    const bool active = tid < n;
    uint64_t* live = nullptr;
    uint32_t live_round = 0;
    for (uint32_t r = 0; r <= rounds; ++r) {
      const bool replace =
          active && (r == 0 || r == rounds || gal_churn_coin(seed, tid, r) != 0);
      void* old = nullptr;
      if (replace && live != nullptr) {
        check_word(live, gal_churn_pattern(tid, live_round), c);
        old = live;
        live = nullptr;
      }
      if (r == rounds) {
        batch::batch_free(s, old, lds);
        break;
      }
      const uint64_t size = gal_churn_size(seed, tid, r);
      uint64_t* np = static_cast<uint64_t*>(batch::batch_step(s, old, replace, size, lds));
      if (replace) {
        c.mallocs += 1;
        if (np) {
          write_word(np, gal_churn_pattern(tid, r), c);
          live = np;
          live_round = r;
        } else {
          c.misses += 1;
        }
      }
    }
  }
  add_counters(ctr, c);
}

template <int kThreads>
int launch_batched_word(gal_allocator* a, uint64_t seed, uint64_t n, uint32_t rounds,
                        uint64_t* ctr, uint32_t workgroups, hipStream_t stream) {
  if (workgroups == 0) {
    int dev = 0, cus = 256, per_cu = 1;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_churn_batched_word<kThreads>,
                                                     kThreads, 0) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    workgroups = (uint32_t)(cus * per_cu);
  }
  const uint64_t chunks = (n + kThreads - 1) / kThreads;
  if (workgroups > chunks) workgroups = (uint32_t)chunks;
  k_churn_batched_word<kThreads><<<workgroups, kThreads, 0, stream>>>(a->dev, seed, n, rounds,
                                                                      ctr);
  return launch_status("churn_batched_word");
}

}  // namespace

extern "C" {

int gal_run_churn_word(gal_allocator* a, uint64_t seed, uint64_t n, uint32_t rounds,
                       uint64_t* ctr, uint32_t batched, uint32_t workgroups, uint32_t threads,
                       void* stream) {
  if (!a || !ctr) return fail(GAL_ERR_INVALID, "null argument");
  if (batched > 1) return fail(GAL_ERR_INVALID, "batched is 0 (per thread) or 1 (workgroups)");
  if (n == 0) return GAL_OK;
  hipStream_t st = (hipStream_t)stream;
  if (!batched) {
    if (threads == 0 || threads > 1024 || threads % 64) threads = 256;
    k_churn_word<<<grid_for(n, threads), threads, 0, st>>>(a->dev, seed, n, rounds, ctr);
    return launch_status("churn_word");
  }
  switch (threads) {
    case 256: return launch_batched_word<256>(a, seed, n, rounds, ctr, workgroups, st);
    case 512: return launch_batched_word<512>(a, seed, n, rounds, ctr, workgroups, st);
    case 0:
    case 1024: return launch_batched_word<1024>(a, seed, n, rounds, ctr, workgroups, st);
    default:
      return fail(GAL_ERR_INVALID, "batched churn runs 256, 512 or 1024 threads per workgroup");
  }
}

}  // extern "C"

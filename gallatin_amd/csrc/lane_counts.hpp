// Per-lane counts of the workload kernels and their per-workgroup sums.
// Included inside an anonymous namespace of each workload source.
#ifndef GALLATIN_AMD_LANE_COUNTS_HPP
#define GALLATIN_AMD_LANE_COUNTS_HPP

// Sum of a per-lane value over the wave (every lane of the wave active).
__device__ inline uint64_t wave_sum(uint64_t v) {
  for (int off = 32; off >= 1; off >>= 1)
    v += wave::fetch(v, (int)(wave::lane() ^ (uint32_t)off));
  return v;
}

// Per-lane counts of one thread, summed over the workgroup (wave sums, then
// LDS) and added with one global atomic per nonzero counter per workgroup.
// The counters share one cache line, where device atomics serialise: one
// atomic per counter per wave took 0.28 ms of a 0.63 ms launch of 2^20
// threads.  Every thread of the workgroup calls it once, at the end.
struct lane_counts {
  uint32_t pairs = 0, misses = 0, bad = 0, mallocs = 0;
  uint64_t bytes = 0;
};

__device__ inline void add_counters(uint64_t* ctr, const lane_counts& c) {
  __shared__ unsigned long long wg[5];
  const uint64_t v[5] = {wave_sum(c.pairs), wave_sum(c.misses), wave_sum(c.bad),
                         wave_sum(c.bytes), wave_sum(c.mallocs)};
  if (threadIdx.x < 5) wg[threadIdx.x] = 0;
  __syncthreads();
  if (wave::lane() == 0) {
#pragma unroll
    for (int i = 0; i < 5; ++i)
      if (v[i]) atomicAdd(&wg[i], (unsigned long long)v[i]);
  }
  __syncthreads();
  if (threadIdx.x < 5 && wg[threadIdx.x])
    utils::atomic_add(&ctr[threadIdx.x], (uint64_t)wg[threadIdx.x]);
}

#endif

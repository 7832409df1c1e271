// A user's translation unit written against the allocator's C++ headers the
// way programs written for the reference use it: device malloc/free inlined
// into the user's own kernels, its own global allocator symbol, and a
// Gallatin<> instance it creates and frees itself.  Reference usage followed:
// tests/src/global_test.cu:33-120 (insert_one_size, free_one_size_pointer)
// and tests/src/gallatin_test.cu:59-112.  user_structures drives the
// allocator-backed data structures from this unit the same way, after
// tests/src/vector_tests.cu:121-137, extendible_ht_test.cu:223-249,
// chaining_test.cu:38-79 and block_queue_test.cu:27-33.
//
// Built into tests/dropin/libuser_kernels.so (linked with libgallatin_hip.so)
// and driven from tests/test_gpu_dropin.py.
#include <hip/hip_runtime.h>

#include "gallatin/allocators/global_allocator.hpp"
#include "gallatin/data_structs/block_queue.hpp"
#include "gallatin/data_structs/chaining_table.hpp"
#include "gallatin/data_structs/extendible_ht.hpp"
#include "gallatin/data_structs/fixed_vector.hpp"
#include "gallatin/data_structs/queue.hpp"

using namespace gallatin::allocators;
using namespace gallatin::data_structs;

namespace {

__global__ void insert_one_size(uint64_t num_inserts, uint64_t size, uint64_t** ptrs,
                                unsigned long long* misses) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid >= num_inserts) return;
  uint64_t* p = static_cast<uint64_t*>(global_malloc(size));
  ptrs[tid] = p;
  if (p == nullptr) {
    atomicAdd(misses, 1ULL);
    return;
  }
  p[0] = tid;
}

__global__ void free_one_size(uint64_t num_inserts, uint64_t** ptrs,
                              unsigned long long* bad) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid >= num_inserts) return;
  uint64_t* p = ptrs[tid];
  if (p == nullptr) return;
  if (p[0] != tid) atomicAdd(bad, 1ULL);
  global_free(p);
}

using my_allocator = Gallatin<16ULL * 1024 * 1024, 16ULL, 4096ULL>;

// alloc_one_size_pointer + free in one kernel on an instance the user owns.
__global__ void own_instance_round(my_allocator* alloc, uint64_t n, uint64_t size,
                                   unsigned long long* counts) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid >= n) return;
  uint32_t* p = static_cast<uint32_t*>(alloc->malloc(size));
  if (p == nullptr) {
    atomicAdd(&counts[0], 1ULL);
    return;
  }
  if (!alloc->owns_allocation(p)) atomicAdd(&counts[1], 1ULL);
  if (alloc->get_allocated_size(size) < size) atomicAdd(&counts[1], 1ULL);
  p[0] = (uint32_t)tid;
  if (p[0] != (uint32_t)tid) atomicAdd(&counts[1], 1ULL);
  alloc->free(p);
}

}  // namespace

// counts_out[0] misses, [1] pattern / ownership errors, for each part:
// [0..1] global allocator, [2..3] own instance.
extern "C" int user_run(uint64_t pool_bytes, uint64_t n, uint64_t size,
                        unsigned long long* counts_out) {
  init_global_allocator(pool_bytes, 42, false);
  uint64_t** ptrs = nullptr;
  unsigned long long* counts = nullptr;
  if (hipMalloc(&ptrs, n * sizeof(uint64_t*)) != hipSuccess) return -1;
  if (hipMalloc(&counts, 4 * sizeof(unsigned long long)) != hipSuccess) return -1;
  if (hipMemset(counts, 0, 4 * sizeof(unsigned long long)) != hipSuccess) return -1;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  insert_one_size<<<blocks, 256>>>(n, size, ptrs, counts);
  free_one_size<<<blocks, 256>>>(n, ptrs, counts + 1);
  if (hipDeviceSynchronize() != hipSuccess) return -2;
  print_global_stats();
  free_global_allocator();

  my_allocator* alloc = my_allocator::generate_on_device(pool_bytes, 7, false);
  if (alloc == nullptr) return -3;
  for (int r = 0; r < 4; ++r) own_instance_round<<<blocks, 256>>>(alloc, n, size, counts + 2);
  if (hipDeviceSynchronize() != hipSuccess) return -2;
  alloc->print_info();
  my_allocator::free_on_device(alloc);

  const hipError_t e = hipMemcpy(counts_out, counts, 4 * sizeof(unsigned long long),
                                 hipMemcpyDeviceToHost);
  (void)hipFree(ptrs);
  (void)hipFree(counts);
  return e == hipSuccess ? 0 : -2;
}

namespace {

// tests/src/global_test_noprint.cu:81-160 with its checks counted instead of
// printed: each thread mallocs, writes its id, and the check kernel finds
// the allocation's segment, tree and block through the table accessors and
// counts allocations per block.
__global__ void table_malloc(uint64_t n, uint64_t size, uint64_t** bitarray) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid >= n) return;
  uint64_t* malloc = static_cast<uint64_t*>(global_malloc(size));
  bitarray[tid] = malloc;
  if (malloc) malloc[0] = tid;
}

__global__ void table_check(uint64_t n, uint64_t size, uint64_t** bitarray,
                            uint64_t* block_counts, unsigned long long* counts) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid >= n) return;
  uint64_t* malloc = bitarray[tid];
  if (malloc == nullptr) {
    atomicAdd(&counts[0], 1ULL);
    return;
  }
  if (tid == 0 && !global_gallatin->check_alloc_valid((void*)malloc)) atomicAdd(&counts[1], 1ULL);
  uint64_t segment = global_gallatin->table->get_segment_from_ptr((void*)malloc);
  uint16_t tree_id = global_gallatin->table->read_tree_id(segment);
  uint64_t offset = global_gallatin->allocation_to_offset((void*)malloc, tree_id);
  uint64_t block_id = offset / 4096;
  atomicAdd((unsigned long long*)&block_counts[block_id], 1ULL);
  // the accessors agree with the geometry: the pointer lies in its segment,
  // the segment belongs to the size's class, the block is that class's
  const uint64_t rel = (uint64_t)((char*)malloc - global_gallatin->memory);
  const uint16_t cls = (uint16_t)(global_gallatin->get_first_bit_bigger(size < 16 ? 16 : size) - 4);
  Block* block_ptr = global_gallatin->table->get_block_from_global_block_id(block_id);
  if ((rel >> 24) != segment || tree_id != cls || block_id / 256 != segment ||
      (block_ptr->malloc_counter >> 20) != tree_id)
    atomicAdd(&counts[1], 1ULL);
  if (malloc[0] != tid) atomicAdd(&counts[2], 1ULL);
}

// check_blocks (global_test_noprint.cu:165-185): no block holds more than
// 4096 allocations.
__global__ void table_check_blocks(const uint64_t* blocks, uint64_t nblocks,
                                   unsigned long long* counts) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid < nblocks && blocks[tid] > 4096) atomicAdd(&counts[3], 1ULL);
}

__global__ void table_free(uint64_t n, uint64_t** bitarray) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid < n && bitarray[tid]) global_free(bitarray[tid]);
}

}  // namespace

// counts_out: [0] misses, [1] accessor disagreements, [2] ids not read
// back, [3] blocks counted past 4096.
extern "C" int user_table_accessors(uint64_t pool_bytes, uint64_t n, uint64_t size,
                                    unsigned long long* counts_out) {
  init_global_allocator(pool_bytes, 42, false);
  const uint64_t nblocks = (pool_bytes >> 24) * 256;
  uint64_t** bitarray = nullptr;
  uint64_t* block_counts = nullptr;
  unsigned long long* counts = nullptr;
  if (hipMalloc(&bitarray, n * sizeof(uint64_t*)) != hipSuccess ||
      hipMalloc(&block_counts, nblocks * sizeof(uint64_t)) != hipSuccess ||
      hipMalloc(&counts, 4 * sizeof(unsigned long long)) != hipSuccess)
    return -1;
  (void)hipMemset(block_counts, 0, nblocks * sizeof(uint64_t));
  (void)hipMemset(counts, 0, 4 * sizeof(unsigned long long));
  const unsigned blocks = (unsigned)((n + 255) / 256);
  table_malloc<<<blocks, 256>>>(n, size, bitarray);
  table_check<<<blocks, 256>>>(n, size, bitarray, block_counts, counts);
  table_check_blocks<<<(unsigned)((nblocks + 255) / 256), 256>>>(block_counts, nblocks, counts);
  table_free<<<blocks, 256>>>(n, bitarray);
  if (hipDeviceSynchronize() != hipSuccess) return -2;
  free_global_allocator();
  const hipError_t e = hipMemcpy(counts_out, counts, 4 * sizeof(unsigned long long),
                                 hipMemcpyDeviceToHost);
  (void)hipFree(bitarray);
  (void)hipFree(block_counts);
  (void)hipFree(counts);
  return e == hipSuccess ? 0 : -2;
}

namespace {

using user_vector = fixed_vector<uint64_t, 16, 1ULL << 20>;
using user_table = extendible_hash_table<uint64_t, 0ULL, ~0ULL, uint64_t, 15, 10, 20>;
using user_chain = chaining_table<uint64_t, uint64_t, 4>;
using user_queue = queue<uint64_t, my_allocator>;
using user_bqueue = block_queue<uint64_t, 256>;

// counts: [0] vector, [1] table, [2] chaining table, [3] queue, [4] block
// queue (failed operations or wrong values); seen: one bit per item

__global__ void vec_insert(user_vector* v, uint64_t n, uint64_t* pos) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid < n) pos[tid] = v->insert(tid);
}

__global__ void vec_check(user_vector* v, uint64_t n, const uint64_t* pos, uint64_t* seen,
                          unsigned long long* counts) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid >= n) return;
  const uint64_t p = pos[tid];
  if (p >= n || (*v)[p] != tid) {
    atomicAdd(&counts[0], 1ULL);
    return;
  }
  const uint64_t bit = 1ULL << (p & 63);
  if (atomicOr((unsigned long long*)&seen[p >> 6], (unsigned long long)bit) & bit)
    atomicAdd(&counts[0], 1ULL);  // two items at one position
}

template <typename T>
__global__ void map_insert(T* t, uint64_t n, unsigned long long* bad) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid < n && !t->insert(tid + 1, tid)) atomicAdd(bad, 1ULL);
}

template <typename T>
__global__ void map_query(T* t, uint64_t n, unsigned long long* bad) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid >= n) return;
  uint64_t v = ~0ULL;
  if (!t->query(tid + 1, v) || v != tid) atomicAdd(bad, 1ULL);
}

// queue<> reports a failed enqueue through try_enqueue; block_queue's
// enqueue returns it (block_queue.cuh:331)
__device__ bool put(user_queue* q, uint64_t x) { return q->try_enqueue(x); }
__device__ bool put(user_bqueue* q, uint64_t x) { return q->enqueue(x); }

template <typename Q>
__global__ void q_enqueue(Q* q, uint64_t n, unsigned long long* bad) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid < n && !put(q, tid)) atomicAdd(bad, 1ULL);
}

template <typename Q>
__global__ void q_dequeue(Q* q, uint64_t n, uint64_t* seen, unsigned long long* bad) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid >= n) return;
  uint64_t v = 0;
  if (!q->dequeue(v) || v >= n) {
    atomicAdd(bad, 1ULL);
    return;
  }
  const uint64_t bit = 1ULL << (v & 63);
  if (atomicOr((unsigned long long*)&seen[v >> 6], (unsigned long long)bit) & bit)
    atomicAdd(bad, 1ULL);  // an item came out twice
}

}  // namespace

// Every structure with n items; counts_out[0..4] as above, all 0 when the
// structures behave.  Needs about 1 GiB of device memory.
extern "C" int user_structures(uint64_t n, unsigned long long* counts_out) {
  init_global_allocator(1ULL << 30, 11, false);
  unsigned long long* counts = nullptr;
  uint64_t *pos = nullptr, *seen = nullptr;
  const uint64_t words = (n + 63) / 64;
  if (hipMalloc(&counts, 5 * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&pos, n * sizeof(uint64_t)) != hipSuccess ||
      hipMalloc(&seen, words * sizeof(uint64_t)) != hipSuccess)
    return -1;
  (void)hipMemset(counts, 0, 5 * sizeof(unsigned long long));
  const unsigned blocks = (unsigned)((n + 255) / 256);

  user_vector* v = user_vector::get_device_vector();
  if (!v) return -2;
  (void)hipMemset(seen, 0, words * sizeof(uint64_t));
  vec_insert<<<blocks, 256>>>(v, n, pos);
  vec_check<<<blocks, 256>>>(v, n, pos, seen, counts);
  if (hipDeviceSynchronize() != hipSuccess) return -3;
  user_vector::free_device_vector(v);

  user_table* t = user_table::generate_on_device(42);
  if (!t) return -4;
  map_insert<<<blocks, 256>>>(t, n, counts + 1);
  map_query<<<blocks, 256>>>(t, n, counts + 1);
  if (hipDeviceSynchronize() != hipSuccess) return -5;
  user_table::free_on_device(t);

  user_chain* c = user_chain::generate_on_device(n, 0, 123456ULL);
  if (!c) return -6;
  map_insert<<<blocks, 256>>>(c, n, counts + 2);
  map_query<<<blocks, 256>>>(c, n, counts + 2);
  if (hipDeviceSynchronize() != hipSuccess) return -7;
  c->print_chain_stats();
  user_chain::free_on_device(c);
  free_global_allocator();

  my_allocator* alloc = my_allocator::generate_on_device(1ULL << 30, 5, false);
  if (!alloc) return -8;
  user_queue* q = user_queue::generate_on_device(alloc);
  if (!q) return -9;
  (void)hipMemset(seen, 0, words * sizeof(uint64_t));
  q_enqueue<<<blocks, 256>>>(q, n, counts + 3);
  q_dequeue<<<blocks, 256>>>(q, n, seen, counts + 3);
  if (hipDeviceSynchronize() != hipSuccess) return -10;
  user_queue::free_on_device(q);
  my_allocator::free_on_device(alloc);

  init_global_allocator(1ULL << 30, 12, false);
  user_bqueue* b = user_bqueue::generate_on_device();
  if (!b) return -11;
  (void)hipMemset(seen, 0, words * sizeof(uint64_t));
  q_enqueue<<<blocks, 256>>>(b, n, counts + 4);
  q_dequeue<<<blocks, 256>>>(b, n, seen, counts + 4);
  if (hipDeviceSynchronize() != hipSuccess) return -12;
  user_bqueue::free_on_device(b);
  free_global_allocator();

  const hipError_t e = hipMemcpy(counts_out, counts, 5 * sizeof(unsigned long long),
                                 hipMemcpyDeviceToHost);
  (void)hipFree(counts);
  (void)hipFree(pos);
  (void)hipFree(seen);
  return e == hipSuccess ? 0 : -13;
}

namespace {

// The reference's churn test kernel (tests/src/gallatin_churn.cu:37-95) as
// a user writes it against these headers: every round a thread hashes its
// running value into a size, mallocs it (retrying up to ten times),
// exchanges its id into the allocation's first word (which must hold 0:
// the pool starts zeroed and every user puts 0 back), exchanges 0 back
// (which must return its id) and frees it.  The reference prints each
// double malloc; here they are counted.
template <typename allocator>
__global__ void pointer_churn(allocator* gallatin, uint64_t num_allocs, int num_rounds,
                              uint64_t min_size, uint64_t max_size,
                              unsigned long long* counts) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid >= num_allocs) return;
  uint64_t hash = tid;
  for (int i = 0; i < num_rounds; ++i) {
    hash = gallatin::hashers::MurmurHash64A(&hash, sizeof(uint64_t), i);
    uint64_t size = hash % max_size;
    if (size < min_size) size = min_size;
    __threadfence();
    void* a = gallatin->malloc(size);
    for (int retry = 0; a == nullptr && retry < 10; ++retry) {
      __threadfence();
      a = gallatin->malloc(size);
    }
    if (a == nullptr) {
      atomicAdd(&counts[0], 1ULL);
      continue;
    }
    auto* word = static_cast<unsigned long long*>(a);
    if (atomicExch(word, (unsigned long long)tid) != 0ULL) atomicAdd(&counts[1], 1ULL);
    if (atomicExch(word, 0ULL) != (unsigned long long)tid) atomicAdd(&counts[2], 1ULL);
    gallatin->free(a);
  }
}

}  // namespace

// tests/src/gallatin_churn.cu:98-165 gallatin_pointer_churn: an instance of
// num_bytes (seed 111), one launch of num_allocs threads in blocks of 512,
// then print_info.  counts_out: [0] misses, [1] double mallocs seen at the
// first exchange, [2] seen at the second, [3] allocator errors after the
// launch.
extern "C" int user_churn(uint64_t num_bytes, uint64_t num_allocs, int num_rounds,
                          uint64_t min_size, uint64_t max_size,
                          unsigned long long* counts_out) {
  unsigned long long* counts = nullptr;
  if (num_allocs == 0 || max_size == 0) return -1;
  if (hipMalloc(&counts, 3 * sizeof(unsigned long long)) != hipSuccess) return -1;
  if (hipMemset(counts, 0, 3 * sizeof(unsigned long long)) != hipSuccess) return -1;
  my_allocator* alloc = my_allocator::generate_on_device(num_bytes, 111, false);
  if (alloc == nullptr) return -3;
  const unsigned blocks = (unsigned)((num_allocs - 1) / 512 + 1);
  pointer_churn<my_allocator><<<blocks, 512>>>(alloc, num_allocs, num_rounds, min_size,
                                                max_size, counts);
  if (hipDeviceSynchronize() != hipSuccess) return -2;
  alloc->print_info();
  gal_info info;
  const int irc = gal_get_info(gal_from_device_state(alloc), &info);
  my_allocator::free_on_device(alloc);
  if (irc != 0) return -4;
  const hipError_t e = hipMemcpy(counts_out, counts, 3 * sizeof(unsigned long long),
                                 hipMemcpyDeviceToHost);
  counts_out[3] = info.errors;
  (void)hipFree(counts);
  return e == hipSuccess ? 0 : -2;
}

namespace {

// tests/src/gallatin_single_thread.cu:38-86: one thread mallocs num_allocs
// allocations of one size in a row through the global allocator, writing
// each index into its first word, and a second launch frees them in order,
// checking the index first.  Counted here instead of printed.
__global__ void single_thread_malloc(uint64_t num_allocs, uint64_t size, uint64_t** ptrs,
                                     unsigned long long* counts) {
  if (gallatin::utils::get_tid() != 0) return;
  for (uint64_t i = 0; i < num_allocs; ++i) {
    uint64_t* p = static_cast<uint64_t*>(global_malloc(size));
    ptrs[i] = p;
    if (p == nullptr) {
      counts[0] += 1;
      continue;
    }
    p[0] = i;
    __threadfence();
  }
}

__global__ void single_thread_free(uint64_t num_allocs, uint64_t** ptrs,
                                   unsigned long long* counts) {
  if (gallatin::utils::get_tid() != 0) return;
  for (uint64_t i = 0; i < num_allocs; ++i) {
    uint64_t* p = ptrs[i];
    if (p == nullptr) continue;
    if (p[0] != i) counts[1] += 1;
    global_free(p);
    __threadfence();
  }
}

}  // namespace

// gallatin_single_thread.cu:91-140: a global allocator of 1000 segments,
// then num_rounds rounds of one single-thread malloc launch and one
// single-thread free launch.  counts_out: [0] misses, [1] indices that did
// not read back, [2] allocator errors, [3] segments not free after the
// rounds and a drain.
extern "C" int user_single_thread(uint64_t num_allocs, int num_rounds, uint64_t size,
                                  unsigned long long* counts_out) {
  if (num_allocs == 0) return -1;
  init_global_allocator(1000ULL * (16ULL << 20), 42, false);
  uint64_t** ptrs = nullptr;
  unsigned long long* counts = nullptr;
  if (hipMalloc(&ptrs, num_allocs * sizeof(uint64_t*)) != hipSuccess) return -1;
  if (hipMalloc(&counts, 2 * sizeof(unsigned long long)) != hipSuccess) return -1;
  if (hipMemset(counts, 0, 2 * sizeof(unsigned long long)) != hipSuccess) return -1;
  for (int r = 0; r < num_rounds; ++r) {
    single_thread_malloc<<<1, 1>>>(num_allocs, size, ptrs, counts);
    single_thread_free<<<1, 1>>>(num_allocs, ptrs, counts);
    if (hipDeviceSynchronize() != hipSuccess) return -2;
  }
  print_global_stats();
  global_allocator_type* g = nullptr;
  if (hipMemcpyFromSymbol(&g, HIP_SYMBOL(global_gallatin), sizeof(g)) != hipSuccess) return -2;
  gal_allocator* h = g ? gal_from_device_state(g) : nullptr;
  gal_info info;
  int rc = h ? gal_drain(h, nullptr) : -4;
  if (rc == 0 && hipDeviceSynchronize() != hipSuccess) rc = -2;
  if (rc == 0) rc = gal_get_info(h, &info);
  free_global_allocator();
  const hipError_t e = hipMemcpy(counts_out, counts, 2 * sizeof(unsigned long long),
                                 hipMemcpyDeviceToHost);
  (void)hipFree(ptrs);
  (void)hipFree(counts);
  if (rc != 0) return -4;
  counts_out[2] = info.errors;
  counts_out[3] = info.num_segments - info.segments_free;
  return e == hipSuccess ? 0 : -2;
}

namespace {

// tests/src/gallatin_host_test.cu:40-185 on a pool in host memory: the
// allocation kernels of that test on an instance the user owns, and its
// overlap check, which finds each allocation's block of the one size class
// from the pool base (alloc->table->memory) and counts allocations per
// block.
__global__ void host_alloc(my_allocator* alloc, uint64_t n, uint64_t size, uint64_t** ptrs,
                           unsigned long long* counts) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid >= n) return;
  uint64_t* p = static_cast<uint64_t*>(alloc->malloc(size));
  ptrs[tid] = p;
  if (p == nullptr) {
    atomicAdd(&counts[0], 1ULL);
    return;
  }
  p[0] = tid;
  __threadfence();
}

__global__ void host_overlaps(my_allocator* alloc, uint64_t slice_size, uint64_t** ptrs,
                              uint64_t n, unsigned long long* per_block, uint64_t nblocks,
                              unsigned long long* counts) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid >= n || ptrs[tid] == nullptr) return;
  const char* start_of_memory = alloc->table->memory;
  const uint64_t offset = ((uint64_t)ptrs[tid] - (uint64_t)start_of_memory) / slice_size;
  if (offset / 4096 < nblocks)
    atomicAdd(&per_block[offset / 4096], 1ULL);
  else
    atomicAdd(&counts[2], 1ULL);  // outside the pool
}

__global__ void host_block_counts(const unsigned long long* per_block, uint64_t nblocks,
                                  unsigned long long* counts) {
  const uint64_t b = gallatin::utils::get_tid();
  if (b >= nblocks) return;
  if (per_block[b] > 4096) atomicAdd(&counts[2], 1ULL);
  if (per_block[b]) atomicAdd(&counts[3], per_block[b]);
}

__global__ void host_free(my_allocator* alloc, uint64_t n, uint64_t** ptrs,
                          unsigned long long* counts) {
  const uint64_t tid = gallatin::utils::get_tid();
  if (tid >= n) return;
  uint64_t* p = ptrs[tid];
  if (p == nullptr) return;
  if (p[0] != tid) atomicAdd(&counts[1], 1ULL);
  alloc->free(p);
  __threadfence();
}

}  // namespace

// counts_out: [0] misses, [1] ids that did not read back, [2] blocks
// holding more than 4096 allocations and allocations outside the pool,
// [3] allocations counted in blocks
// (n - misses when none is lost), [4] allocator errors.  size must be a
// slice class (a power of two in [16, 4096]).
extern "C" int user_host_pool(uint64_t pool_bytes, uint64_t n, uint64_t size,
                              unsigned long long* counts_out) {
  if (n == 0 || size < 16 || size > 4096 || (size & (size - 1))) return -1;
  my_allocator* alloc = my_allocator::generate_on_device_host(pool_bytes, 42, false);
  if (alloc == nullptr) return -3;
  gal_info info;
  if (gal_get_info(gal_from_device_state(alloc), &info) != 0) return -4;
  const uint64_t nblocks = info.num_segments * (info.bytes_per_segment / (size * 4096));
  uint64_t** ptrs = nullptr;
  unsigned long long *counts = nullptr, *per_block = nullptr;
  if (hipMalloc(&ptrs, n * sizeof(uint64_t*)) != hipSuccess) return -1;
  if (hipMalloc(&counts, 4 * sizeof(unsigned long long)) != hipSuccess) return -1;
  if (hipMalloc(&per_block, nblocks * sizeof(unsigned long long)) != hipSuccess) return -1;
  if (hipMemset(counts, 0, 4 * sizeof(unsigned long long)) != hipSuccess) return -1;
  if (hipMemset(per_block, 0, nblocks * sizeof(unsigned long long)) != hipSuccess) return -1;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  host_alloc<<<blocks, 256>>>(alloc, n, size, ptrs, counts);
  host_overlaps<<<blocks, 256>>>(alloc, size, ptrs, n, per_block, nblocks, counts);
  host_block_counts<<<(unsigned)((nblocks + 255) / 256), 256>>>(per_block, nblocks, counts);
  host_free<<<blocks, 256>>>(alloc, n, ptrs, counts);
  if (hipDeviceSynchronize() != hipSuccess) return -2;
  alloc->print_info();
  const int irc = gal_get_info(gal_from_device_state(alloc), &info);
  my_allocator::free_on_device(alloc);
  const hipError_t e = hipMemcpy(counts_out, counts, 4 * sizeof(unsigned long long),
                                 hipMemcpyDeviceToHost);
  (void)hipFree(ptrs);
  (void)hipFree(counts);
  (void)hipFree(per_block);
  if (irc != 0) return -4;
  counts_out[4] = info.errors;
  return e == hipSuccess ? 0 : -2;
}

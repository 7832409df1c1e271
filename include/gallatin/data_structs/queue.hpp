// Allocator-backed multi-producer multi-consumer FIFO queue (queue.cuh:59-192)
// and the ring of queues built on it (queue.cuh:196-253).
//
// Nodes come from the backing allocator (`allocator->malloc` / `->free`, as
// in the reference).  The list always starts with a dummy node (the
// Michael-Scott queue): enqueue links a node after the last one and swings
// `tail`; dequeue reads the item of the node after the dummy, swings `head`
// to it (which becomes the dummy) and frees the old dummy at once.
//
// The reference links nodes with plain pointers and frees dequeued nodes at
// once, so a dequeuer that read `head` before another dequeuer freed and
// re-linked the same node can swing `head` to a node that is gone (ABA),
// and a queue emptied to one node can have `tail` left on a freed node.
// Here `head`, `tail` and every `next` are tagged words, (count << 48) |
// address, and each update bumps the count, so a stale compare-and-swap
// fails.  A node another dequeuer freed may still be read before such a
// compare-and-swap fails; its memory stays in the allocator's pool (it is
// never unmapped), and an address read from it is used only if it lies in
// the pool.  Node words are read and written with device-coherent atomics,
// since producers and consumers may run on different XCDs.  Every retry
// loop is bounded; enqueue reports failure (no node, or no progress).
#ifndef GALLATIN_QUEUE_HPP
#define GALLATIN_QUEUE_HPP

#include <hip/hip_runtime.h>

#include "gallatin/allocators/alloc_utils.hpp"

namespace gallatin {
namespace data_structs {

namespace tagged {
constexpr uint64_t kAddrMask = (1ULL << 48) - 1;
__device__ __forceinline__ uint64_t addr(uint64_t w) { return w & kAddrMask; }
__device__ __forceinline__ uint64_t count(uint64_t w) { return w >> 48; }
__device__ __forceinline__ uint64_t make(uint64_t a, uint64_t c) {
  return (a & kAddrMask) | (c << 48);
}
}  // namespace tagged

// queue.cuh:19-43.  `item` is kept in 8-byte words.
template <typename T>
struct queue_node {
  static constexpr int kWords = (int)((sizeof(T) + 7) / 8);
  uint64_t next;
  uint64_t words[kWords];

  __device__ void store_item(const T& v) {
    uint64_t tmp[kWords] = {};
    __builtin_memcpy(tmp, &v, sizeof(T));
    for (int i = 0; i < kWords; ++i) gallatin::utils::atomic_store(&words[i], tmp[i]);
  }
  __device__ T load_item() {
    uint64_t tmp[kWords];
    for (int i = 0; i < kWords; ++i) tmp[i] = gallatin::utils::atomic_load(&words[i]);
    T v;
    __builtin_memcpy(&v, tmp, sizeof(T));
    return v;
  }
};

template <typename queue_type, typename allocator>
__global__ void queue_init_kernel(queue_type* q, allocator* a, uint32_t* ok) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *ok = q->init(a) ? 1u : 0u;
}

template <typename queue_type>
__global__ void queue_teardown_kernel(queue_type* q) {
  if (blockIdx.x == 0 && threadIdx.x == 0) q->teardown();
}

template <typename T, typename allocator>
struct queue {
  using my_type = queue<T, allocator>;
  using node = queue_node<T>;
  static constexpr uint32_t kTries = 1u << 20;

  uint64_t head;  // tagged
  uint64_t tail;  // tagged
  allocator* my_backing_allocator;
  uint64_t pool_lo, pool_hi;  // the allocator's pool, for address checks

  // queue.cuh:72-81.  Returns nullptr on failure.
  static __host__ my_type* generate_on_device(allocator* backing_allocator) {
    my_type* dev = nullptr;
    uint32_t* ok = nullptr;
    if (hipMalloc(&dev, sizeof(my_type)) != hipSuccess) return nullptr;
    if (hipMalloc(&ok, sizeof(uint32_t)) != hipSuccess) {
      (void)hipFree(dev);
      return nullptr;
    }
    queue_init_kernel<my_type, allocator><<<1, 1>>>(dev, backing_allocator, ok);
    uint32_t h = 0;
    const bool good = hipMemcpy(&h, ok, sizeof(h), hipMemcpyDeviceToHost) == hipSuccess && h;
    (void)hipFree(ok);
    if (!good) {
      (void)hipFree(dev);
      return nullptr;
    }
    return dev;
  }

  // Frees every node left in the queue and the queue object.
  static __host__ void free_on_device(my_type* dev) {
    if (!dev) return;
    queue_teardown_kernel<my_type><<<1, 1>>>(dev);
    (void)hipDeviceSynchronize();
    (void)hipFree(dev);
  }

  // queue.cuh:83-88, with the dummy node.
  __device__ bool init(allocator* a) {
    my_backing_allocator = a;
    pool_lo = (uint64_t)a->memory;
    pool_hi = pool_lo + (a->num_segments << a->seg_shift);
    node* d = static_cast<node*>(a->malloc(sizeof(node)));
    if (!d) return false;
    gallatin::utils::atomic_store(&d->next, (uint64_t)0);
    __builtin_amdgcn_s_waitcnt(0);
    head = tail = tagged::make((uint64_t)d, 0);
    return true;
  }

  __device__ void teardown() {
    uint64_t h = tagged::addr(head);
    while (h) {
      node* n = reinterpret_cast<node*>(h);
      h = tagged::addr(n->next);
      my_backing_allocator->free(n);
    }
    head = tail = 0;
  }

  __device__ __forceinline__ bool in_pool(uint64_t a) const {
    return a >= pool_lo && a < pool_hi;
  }

  // queue.cuh:90-135.  False when no node could be allocated or no progress
  // was made within the retry bound.
  __device__ bool try_enqueue(T item) {
    node* nd = static_cast<node*>(my_backing_allocator->malloc(sizeof(node)));
    if (!nd) return false;
    nd->store_item(item);
    gallatin::utils::atomic_store(&nd->next, (uint64_t)0);
    __builtin_amdgcn_s_waitcnt(0);  // the node is complete before it is linked
    for (uint32_t tries = 0; tries < kTries; ++tries) {
      const uint64_t t = gallatin::utils::atomic_load(&tail);
      node* tn = reinterpret_cast<node*>(tagged::addr(t));
      const uint64_t n = gallatin::utils::atomic_load(&tn->next);
      if (t != gallatin::utils::atomic_load(&tail)) continue;
      if (tagged::addr(n) == 0) {
        const uint64_t linked = tagged::make((uint64_t)nd, tagged::count(n) + 1);
        if (gallatin::utils::atomic_cas(&tn->next, n, linked) == n) {
          gallatin::utils::atomic_cas(&tail, t, tagged::make((uint64_t)nd, tagged::count(t) + 1));
          return true;
        }
      } else if (in_pool(tagged::addr(n))) {  // tail is behind: help it along
        gallatin::utils::atomic_cas(&tail, t, tagged::make(tagged::addr(n), tagged::count(t) + 1));
      }
      __builtin_amdgcn_s_sleep(1);
    }
    my_backing_allocator->free(nd);
    return false;
  }

  __device__ void enqueue(T item) { (void)try_enqueue(item); }

  // queue.cuh:139-187.  False when the queue is empty.
  __device__ bool dequeue(T& return_val) {
    for (uint32_t tries = 0; tries < kTries; ++tries) {
      const uint64_t h = gallatin::utils::atomic_load(&head);
      const uint64_t t = gallatin::utils::atomic_load(&tail);
      node* hn = reinterpret_cast<node*>(tagged::addr(h));
      const uint64_t n = gallatin::utils::atomic_load(&hn->next);
      if (h != gallatin::utils::atomic_load(&head)) continue;
      const uint64_t na = tagged::addr(n);
      if (tagged::addr(h) == tagged::addr(t)) {
        if (na == 0) return false;  // empty
        if (in_pool(na))
          gallatin::utils::atomic_cas(&tail, t, tagged::make(na, tagged::count(t) + 1));
        continue;
      }
      if (!in_pool(na)) continue;  // read from a node already recycled
      const T v = reinterpret_cast<node*>(na)->load_item();
      if (gallatin::utils::atomic_cas(&head, h, tagged::make(na, tagged::count(h) + 1)) == h) {
        return_val = v;
        my_backing_allocator->free(hn);
        return true;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    return false;
  }
};

template <typename queue_type>
__global__ void ring_init_kernel(queue_type* q, typename queue_type::alloc_type* a,
                                 uint32_t* ok) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *ok = q->init(a) ? 1u : 0u;
}

// queue.cuh:196-253: `width` queues used round robin, one counter for
// enqueues and one for dequeues, to spread the contention of one queue.
template <typename T, int width, typename allocator>
struct circular_CAS_queue {
  using my_type = circular_CAS_queue<T, width, allocator>;
  using alloc_type = allocator;

  uint64_t malloc_counter;
  uint64_t free_counter;
  queue<T, allocator> queue_list[width];

  __device__ bool init(allocator* a) {
    malloc_counter = 0;
    free_counter = 0;
    for (int i = 0; i < width; ++i)
      if (!queue_list[i].init(a)) return false;
    return true;
  }

  static __host__ my_type* generate_on_device(allocator* a) {
    my_type* dev = nullptr;
    uint32_t* ok = nullptr;
    if (hipMalloc(&dev, sizeof(my_type)) != hipSuccess) return nullptr;
    if (hipMalloc(&ok, sizeof(uint32_t)) != hipSuccess) {
      (void)hipFree(dev);
      return nullptr;
    }
    ring_init_kernel<my_type><<<1, 1>>>(dev, a, ok);
    uint32_t h = 0;
    const bool good = hipMemcpy(&h, ok, sizeof(h), hipMemcpyDeviceToHost) == hipSuccess && h;
    (void)hipFree(ok);
    if (!good) {
      (void)hipFree(dev);
      return nullptr;
    }
    return dev;
  }

  __device__ void teardown() {
    for (int i = 0; i < width; ++i) queue_list[i].teardown();
  }

  static __host__ void free_on_device(my_type* dev) {
    if (!dev) return;
    queue_teardown_kernel<my_type><<<1, 1>>>(dev);
    (void)hipDeviceSynchronize();
    (void)hipFree(dev);
  }

  __device__ bool try_enqueue(T item) {
    const uint64_t i = gallatin::utils::atomic_add(&malloc_counter, (uint64_t)1) % width;
    return queue_list[i].try_enqueue(item);
  }
  __device__ void enqueue(T item) { (void)try_enqueue(item); }

  // As in the reference, a dequeue looks at one queue only, so it can find
  // that queue empty while others hold items.
  __device__ bool dequeue(T& return_val) {
    const uint64_t i = gallatin::utils::atomic_add(&free_counter, (uint64_t)1) % width;
    return queue_list[i].dequeue(return_val);
  }
};

}  // namespace data_structs
}  // namespace gallatin

#endif

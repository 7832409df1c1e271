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
// never unmapped), and an address read from it is used only if it could be
// a node of the pool.  Node words are read and written with device-coherent
// atomics, since producers and consumers may run on different XCDs.  Every
// retry loop is bounded; enqueue reports failure (no node, or no progress).
//
// Lanes of one wave that work on the same queue act as one: their enqueues
// append one chain of nodes with one compare-and-swap, and their dequeues
// move `head` past as many nodes with one compare-and-swap, where the
// reference makes every lane compete for the same word.
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

  // An address that may be a node: inside the pool, with room for a whole
  // node, and aligned as every slice is (16 bytes).
  __device__ __forceinline__ bool in_pool(uint64_t a) const {
    return a >= pool_lo && a <= pool_hi - sizeof(node) && (a & 15) == 0;
  }

  // queue.cuh:90-135.  False when no node could be allocated or no progress
  // was made within the retry bound.  The lanes of a wave that enqueue into
  // this queue together link their nodes into one chain in lane order, and
  // one lane appends the chain with a single compare-and-swap.
  __device__ bool try_enqueue(T item) {
    node* nd = static_cast<node*>(my_backing_allocator->malloc(sizeof(node)));
    if (nd) {
      nd->store_item(item);
      gallatin::utils::atomic_store(&nd->next, (uint64_t)0);
    }
    const bool have = nd != nullptr;
    const uint64_t group = gallatin::wave::match(__ballot(have), have, (uint64_t)this);
    if (!have) return false;
    const uint32_t lane = gallatin::wave::lane();
    const uint64_t above = group & ~((2ULL << lane) - 1);  // group lanes after this one
    const int leader = gallatin::wave::first(group);
    const int last = 63 - __builtin_clzll(group);
    const uint64_t succ = gallatin::wave::fetch((uint64_t)nd, above ? gallatin::wave::first(above) : (int)lane);
    if (above) gallatin::utils::atomic_store(&nd->next, tagged::make(succ, 0));
    __builtin_amdgcn_s_waitcnt(0);  // every node of the chain is complete before it is linked
    const uint64_t first_node = gallatin::wave::fetch((uint64_t)nd, leader);
    const uint64_t last_node = gallatin::wave::fetch((uint64_t)nd, last);
    uint32_t linked = 0;
    if ((int)lane == leader) {
      for (uint32_t tries = 0; tries < kTries; ++tries) {
        const uint64_t t = gallatin::utils::atomic_load(&tail);
        node* tn = reinterpret_cast<node*>(tagged::addr(t));
        const uint64_t n = gallatin::utils::atomic_load(&tn->next);
        if (t != gallatin::utils::atomic_load(&tail)) continue;
        if (tagged::addr(n) == 0) {
          const uint64_t link = tagged::make(first_node, tagged::count(n) + 1);
          if (gallatin::utils::atomic_cas(&tn->next, n, link) == n) {
            gallatin::utils::atomic_cas(&tail, t, tagged::make(last_node, tagged::count(t) + 1));
            linked = 1;
            break;
          }
        } else if (in_pool(tagged::addr(n))) {  // tail is behind: help it along
          gallatin::utils::atomic_cas(&tail, t,
                                      tagged::make(tagged::addr(n), tagged::count(t) + 1));
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    linked = gallatin::wave::fetch(linked, leader);
    if (!linked) my_backing_allocator->free(nd);
    return linked != 0;
  }

  __device__ void enqueue(T item) { (void)try_enqueue(item); }

  // queue.cuh:139-187.  False when the queue is empty.  The lanes of a wave
  // that dequeue from this queue together walk up to that many nodes from
  // `head` in step, read the items of those nodes, then move `head` past
  // all of them with one compare-and-swap; lane r keeps the item of the
  // r-th node, and the nodes that stop being the dummy are freed.  Items
  // are read before the compare-and-swap, as in the one-node form: the last
  // node taken becomes the dummy, which a later dequeue may free at once.
  __device__ bool dequeue(T& return_val) {
    const uint64_t group = gallatin::wave::match(__ballot(1), true, (uint64_t)this);
    const uint32_t lane = gallatin::wave::lane();
    const int leader = gallatin::wave::first(group);
    const uint32_t k = gallatin::wave::count(group);
    const uint32_t rank = gallatin::wave::count(group & gallatin::wave::lanes_below());
    // A wave may hold groups of several queues: each group runs the loop
    // below with values broadcast from its leader, so every branch is
    // uniform within the group.
    uint64_t mine = 0;   // the node whose item this lane takes
    uint32_t got = 0;    // nodes taken by the group (0 = empty)
    uint64_t old_head = 0;
    T val{};
    bool done = false;
    for (uint32_t tries = 0; tries < kTries && !done; ++tries) {
      const uint64_t h = gallatin::wave::fetch(gallatin::utils::atomic_load(&head), leader);
      const uint64_t t = gallatin::wave::fetch(gallatin::utils::atomic_load(&tail), leader);
      // walk up to k nodes after the dummy
      uint64_t cur = tagged::addr(h);
      uint32_t m = 0;
      bool stale = false;
      mine = 0;
      for (uint32_t i = 0; i < k; ++i) {
        const uint64_t nx = gallatin::wave::fetch(
            gallatin::utils::atomic_load(&reinterpret_cast<node*>(cur)->next), leader);
        const uint64_t na = tagged::addr(nx);
        if (na == 0) break;
        if (!in_pool(na)) {  // read from a node already recycled
          stale = true;
          break;
        }
        if (rank == i) mine = na;
        cur = na;
        m = i + 1;
      }
      const uint64_t h2 = gallatin::wave::fetch(gallatin::utils::atomic_load(&head), leader);
      if (stale || h2 != h) continue;
      if (m == 0) {
        if (tagged::addr(h) == tagged::addr(t)) {  // empty
          done = true;
          break;
        }
        continue;
      }
      // tail must not stay on a node about to be freed: the dummy or any
      // taken node but the last (which becomes the dummy)
      const uint64_t ta = tagged::addr(t);
      const bool tail_in =
          ta == tagged::addr(h) || (__ballot(rank + 1 < m && mine == ta) & group) != 0;
      if (tail_in) {
        if ((int)lane == leader) {
          const uint64_t tn = gallatin::utils::atomic_load(&reinterpret_cast<node*>(ta)->next);
          if (in_pool(tagged::addr(tn)))
            gallatin::utils::atomic_cas(&tail, t,
                                        tagged::make(tagged::addr(tn), tagged::count(t) + 1));
        }
        continue;
      }
      if (rank < m) val = reinterpret_cast<node*>(mine)->load_item();
      uint32_t won = 0;
      if ((int)lane == leader)
        won = gallatin::utils::atomic_cas(&head, h, tagged::make(cur, tagged::count(h) + 1)) == h;
      won = gallatin::wave::fetch(won, leader);
      if (won) {
        got = m;
        old_head = tagged::addr(h);
        done = true;
      } else {
        __builtin_amdgcn_s_sleep(1);
      }
    }
    const bool take = rank < got;
    if (take) return_val = val;
    // the node before each taken one stops being the dummy: lane 0 of the
    // group frees the old dummy, lane r frees the node taken by lane r - 1
    const uint64_t below = group & gallatin::wave::lanes_below();
    const int prev = below ? 63 - __builtin_clzll(below) : (int)lane;
    const uint64_t prev_node = gallatin::wave::fetch(mine, prev);
    __builtin_amdgcn_s_waitcnt(0);  // items are read before their nodes are freed
    if (take) my_backing_allocator->free(reinterpret_cast<node*>(rank == 0 ? old_head : prev_node));
    return take;
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

// Allocator-backed queue of blocks of items (block_queue.cuh:25-463).
//
// The queue is a list of nodes taken from the global allocator, each with
// room for items_per_block items.  enqueue claims the next position of the
// last node with one atomic (one per wave for the lanes on the same node),
// writes the item there and sets the position's bit in the node's `ready`
// mask.  When the node is full, the lane that swaps the node's `next` from
// 0 to a marker allocates and links a new node, and every lane tries again
// there.  The reference's enqueue claims, writes and links the same way
// (block_queue_node::enqueue, block_queue.cuh:86-100; block_queue::enqueue,
// :331-405), except that there every lane that finds the node full
// allocates a node and the losers of the link keep theirs.
//
// The reference has no working dequeue (block_queue.cuh:409-457 is
// commented out, and the node's grab_active, :107-125, adds 1 to a 64-bit
// word at the address of a 32-bit counter).  Here dequeue takes the first
// node that still holds items: it claims positions d, d + 1, ... of that node
// with a compare-and-swap of the node's `taken` counter, one for the lanes
// of a wave on that node and only as many as are ready in a row from d, so
// a dequeuer never holds a claim while it waits and items leave a node in
// position order.  A position that is claimed but not yet written makes
// dequeue report the queue empty for now.  When every position of
// the first node is taken, `head` moves to the next node and the old node
// goes on a retired list.  Retired nodes are returned to the allocator by
// reclaim() and free_on_device(), which need a queue no kernel is using: a
// thread may still read, or compare-and-swap, a node another thread retired.
//
// Order: items of one node leave in the order their positions were
// claimed, and nodes leave in list order, so one producer's items come out
// in the order it enqueued them.  Every loop is bounded; enqueue reports
// failure when no node could be allocated.  Node words are read and written
// with device-coherent atomics.
#ifndef GALLATIN_BLOCK_QUEUE_HPP
#define GALLATIN_BLOCK_QUEUE_HPP

#include <hip/hip_runtime.h>

#include "gallatin/allocators/global_allocator.hpp"

namespace gallatin {
namespace data_structs {

// block_queue.cuh:25-191.
template <typename T, int num_items>
struct block_queue_node {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit items");
  static_assert(num_items >= 1 && num_items <= (1 << 16), "items per node");
  static constexpr int kReadyWords = (num_items + 63) / 64;

  uint64_t next;     // address of the next node, 0 = none
  uint64_t retired;  // next node on the retired list
  uint32_t want;     // positions handed to enqueuers (may pass num_items)
  uint32_t taken;    // positions claimed by dequeuers
  uint64_t ready[kReadyWords];  // bit p: the item at position p is written
  T items[num_items];

  __device__ void init() {
    gallatin::utils::atomic_store(&next, (uint64_t)0);
    gallatin::utils::atomic_store(&retired, (uint64_t)0);
    gallatin::utils::atomic_store(&want, 0u);
    gallatin::utils::atomic_store(&taken, 0u);
    for (int w = 0; w < kReadyWords; ++w) gallatin::utils::atomic_store(&ready[w], (uint64_t)0);
  }

  // block_queue.cuh:86-100, after the position is claimed: the item, then
  // its ready bit.
  __device__ void place(uint32_t p, T item) {
    gallatin::utils::atomic_store(&items[p], item);
    __builtin_amdgcn_s_waitcnt(0);  // the item lands before its ready bit
    gallatin::utils::atomic_or(&ready[p >> 6], (uint64_t)1 << (p & 63));
  }
};

template <typename queue>
__global__ void block_queue_init(queue* q, uint32_t* ok) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *ok = q->init() ? 1u : 0u;
}

template <typename queue>
__global__ void block_queue_reclaim(queue* q, uint64_t* freed, int all) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *freed = q->reclaim(all != 0);
}

template <typename T, int items_per_block>
struct block_queue {
  using my_type = block_queue<T, items_per_block>;
  using node_type = block_queue_node<T, items_per_block>;
  static constexpr uint32_t kTries = 1u << 20;  // attempts per operation
  static constexpr uint32_t kHops = 1u << 24;   // bound on a walk along the list

  uint64_t head;      // first node that may hold items
  uint64_t tail;      // a node at or before the last one
  uint64_t retired;   // retired nodes, linked by their `retired` word
  uint64_t failed;    // enqueues refused for want of a node

  // ---- host ----
  // block_queue.cuh:255-272.  Needs the global allocator.
  static __host__ my_type* generate_on_device() {
    my_type* dev = nullptr;
    uint32_t* ok = nullptr;
    if (hipMalloc(&dev, sizeof(my_type)) != hipSuccess) return nullptr;
    if (hipMalloc(&ok, sizeof(uint32_t)) != hipSuccess) {
      (void)hipFree(dev);
      return nullptr;
    }
    block_queue_init<my_type><<<1, 1>>>(dev, ok);
    uint32_t h = 0;
    const bool good = hipMemcpy(&h, ok, sizeof(h), hipMemcpyDeviceToHost) == hipSuccess && h;
    (void)hipFree(ok);
    if (!good) {
      (void)hipFree(dev);
      return nullptr;
    }
    return dev;
  }

  // Returns retired nodes (all = false) or every node (all = true) to the
  // allocator; the number of nodes freed.  No kernel may use the queue.
  static __host__ uint64_t reclaim_on_device(my_type* dev, bool all) {
    uint64_t* d = nullptr;
    uint64_t n = 0;
    if (!dev || hipMalloc(&d, sizeof(uint64_t)) != hipSuccess) return 0;
    block_queue_reclaim<my_type><<<1, 1>>>(dev, d, all ? 1 : 0);
    (void)hipMemcpy(&n, d, sizeof(n), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return n;
  }

  static __host__ void free_on_device(my_type* dev) {
    if (!dev) return;
    (void)reclaim_on_device(dev, true);
    (void)hipFree(dev);
  }

  // ---- device ----
  static __device__ __forceinline__ uint64_t ld(const uint64_t* p) {
    return gallatin::utils::atomic_load(p);
  }
  static __device__ __forceinline__ node_type* as_node(uint64_t a) {
    return reinterpret_cast<node_type*>(a);
  }

  // block_queue.cuh:276-283, with the dummy-free start: one empty node.
  __device__ bool init() {
    node_type* n =
        static_cast<node_type*>(gallatin::allocators::global_malloc(sizeof(node_type)));
    if (!n) return false;
    n->init();
    __builtin_amdgcn_s_waitcnt(0);
    head = tail = reinterpret_cast<uint64_t>(n);
    retired = 0;
    failed = 0;
    return true;
  }

  // `next` while one thread allocates the node that follows.
  static constexpr uint64_t kAppending = 1;

  // The last node, following links from t.
  __device__ uint64_t last_from(uint64_t t) {
    for (uint32_t h = 0; h < kHops; ++h) {
      const uint64_t n = ld(&as_node(t)->next);
      if (n <= kAppending) break;
      t = n;
    }
    return t;
  }

  // block_queue.cuh:331-405.  Lanes of a wave that reach the same last node
  // claim their positions with one atomic.  When the node is full, the lane
  // that swaps its `next` from 0 to kAppending allocates the next node in
  // the same pass of the retry loop, and the others try again; no lane
  // waits on another inside one pass, so lanes of one wave cannot block
  // each other.  False when no node could be allocated.
  __device__ bool enqueue(T new_item) {
    uint64_t t = ld(&tail);
    int state = 0;  // 0 pending, 1 stored, 2 refused
    for (uint32_t tries = 0; tries < kTries && __ballot(state == 0); ++tries) {
      const bool pending = state == 0;
      if (pending) t = last_from(t);
      const uint64_t group = gallatin::wave::match(__ballot(pending), pending, t);
      if (!pending) continue;
      node_type* nd = as_node(t);
      const int leader = gallatin::wave::first(group);
      uint32_t base = 0;
      if ((int)gallatin::wave::lane() == leader)
        base = gallatin::utils::atomic_add(&nd->want, gallatin::wave::count(group));
      base = gallatin::wave::fetch(base, leader);
      const uint32_t p = base + gallatin::wave::count(group & gallatin::wave::lanes_below());
      if (p < (uint32_t)items_per_block) {
        nd->place(p, new_item);
        state = 1;
        continue;
      }
      // the node is full: one lane appends, the others look again
      uint64_t nx = ld(&nd->next);
      if (nx == 0 &&
          gallatin::utils::atomic_cas(&nd->next, (uint64_t)0, kAppending) == 0) {
        node_type* n =
            static_cast<node_type*>(gallatin::allocators::global_malloc(sizeof(node_type)));
        if (n) {
          n->init();
          __builtin_amdgcn_s_waitcnt(0);  // formatted before it is linked
          gallatin::utils::atomic_store(&nd->next, reinterpret_cast<uint64_t>(n));
        } else {
          gallatin::utils::atomic_store(&nd->next, (uint64_t)0);
          state = 2;
        }
      } else if (nx == kAppending) {
        __builtin_amdgcn_s_sleep(2);
      }
      nx = ld(&nd->next);
      if (nx > kAppending) {
        if (ld(&tail) == t) gallatin::utils::atomic_cas(&tail, t, nx);
        t = nx;
      }
    }
    if (state != 1) gallatin::utils::atomic_add(&failed, (uint64_t)1);
    return state == 1;
  }

  // Ready positions of node n in a row from d, at most k.
  static __device__ uint32_t ready_run(node_type* n, uint32_t d, uint32_t k) {
    uint32_t run = 0;
    while (run < k && d + run < (uint32_t)items_per_block) {
      const uint32_t q = d + run;
      const uint64_t bits = gallatin::utils::atomic_load(&n->ready[q >> 6]) >> (q & 63);
      const uint32_t ones = (uint32_t)gallatin::utils::cfcll(bits);  // trailing ones
      const uint32_t span = 64 - (q & 63);
      run += ones < span ? ones : span;
      if (ones < span) break;
    }
    return run < k ? run : k;
  }

  // False when the queue is empty (or its first item is still being
  // written).  Lanes of a wave on the same first node take their positions
  // with one compare-and-swap, as many as are ready in a row there.
  __device__ bool dequeue(T& return_val) {
    int state = 0;  // 0 pending, 1 got an item, 2 empty
    for (uint32_t tries = 0; tries < kTries && __ballot(state == 0); ++tries) {
      const bool pending = state == 0;
      const uint64_t h = pending ? ld(&head) : 0;
      const uint64_t group = gallatin::wave::match(__ballot(pending), pending, h);
      if (!pending) continue;
      node_type* n = as_node(h);
      const int leader = gallatin::wave::first(group);
      // leader's verdict: bit 63 empty, else (d << 32) | m with m taken
      // (m = 0: look again)
      uint64_t res = 0;
      if ((int)gallatin::wave::lane() == leader) {
        const uint32_t d = gallatin::utils::atomic_load(&n->taken);
        if (d < (uint32_t)items_per_block) {
          const uint32_t m = ready_run(n, d, gallatin::wave::count(group));
          if (m == 0)
            res = 1ULL << 63;  // nothing written at the front
          else if (gallatin::utils::atomic_cas(&n->taken, d, d + m) == d)
            res = ((uint64_t)d << 32) | m;
        } else {  // every position is taken: move to the next node
          const uint64_t nx = ld(&n->next);
          if (nx <= kAppending)
            res = 1ULL << 63;
          else if (gallatin::utils::atomic_cas(&head, h, nx) == h)
            retire(n);
        }
      }
      res = gallatin::wave::fetch(res, leader);
      if (res >> 63) {
        state = 2;
        continue;
      }
      const uint32_t rank = gallatin::wave::count(group & gallatin::wave::lanes_below());
      const uint32_t m = (uint32_t)res;
      if (rank < m) {
        return_val = gallatin::utils::atomic_load(&n->items[(uint32_t)(res >> 32) + rank]);
        state = 1;
      }
    }
    return state == 1;
  }

  // Pushes a node that left the list onto the retired list.
  __device__ void retire(node_type* n) {
    uint64_t top = ld(&retired);
    for (uint32_t tries = 0; tries < kTries; ++tries) {
      gallatin::utils::atomic_store(&n->retired, top);
      __builtin_amdgcn_s_waitcnt(0);
      const uint64_t seen =
          gallatin::utils::atomic_cas(&retired, top, reinterpret_cast<uint64_t>(n));
      if (seen == top) return;
      top = seen;
    }
  }

  // One thread, no concurrent users.
  __device__ uint64_t reclaim(bool all) {
    uint64_t freed = 0;
    uint64_t r = retired;
    retired = 0;
    for (uint32_t h = 0; r && h < kHops; ++h) {
      node_type* n = as_node(r);
      r = n->retired;
      gallatin::allocators::global_free(n);
      ++freed;
    }
    if (all) {
      uint64_t a = head;
      head = tail = 0;
      for (uint32_t h = 0; a > kAppending && h < kHops; ++h) {
        node_type* n = as_node(a);
        a = n->next;
        gallatin::allocators::global_free(n);
        ++freed;
      }
    } else {
      // the tail hint may name a node just freed: point it at the last node
      uint64_t t = head;
      for (uint32_t h = 0; h < kHops && as_node(t)->next > kAppending; ++h) t = as_node(t)->next;
      tail = t;
    }
    return freed;
  }
};

}  // namespace data_structs
}  // namespace gallatin

#endif

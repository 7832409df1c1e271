"""Allocator-backed queues on the GPU.  The reference's queue test is stale
(tests/src/queue_test.cu), so the checks are invariants: every item
enqueued is dequeued exactly once, a single producer's items come out in
order, an empty queue reports empty, producers and consumers may run at the
same time, and after the queue is destroyed the allocator drains to its
fresh state (no node leaked)."""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu
S = 16 << 20


def _alloc(gpu):
    return gpu.Gallatin(S, 16, 4096).generate_on_device(4 << 30, 21, print_info=False)


def _done(a):
    a.drain()
    H.check_fresh(a)
    a.free_on_device()


def test_queue_parallel_roundtrip(gpu):
    import torch
    a = _alloc(gpu)
    q = gpu.Queue(a)
    n = 1 << 20
    items = torch.arange(1, n + 1, dtype=torch.int64, device="cuda")
    ok = q.enqueue(items)
    out, got = q.dequeue(n)
    torch.cuda.synchronize()
    assert bool(ok.all()) and bool(got.all())
    assert np.array_equal(np.sort(out.cpu().numpy()), np.arange(1, n + 1))
    _, more = q.dequeue(64)
    torch.cuda.synchronize()
    assert not bool(more.any()), "an empty queue returned items"
    q.destroy()
    _done(a)


def test_queue_single_producer_fifo(gpu):
    import torch
    a = _alloc(gpu)
    q = gpu.Queue(a)
    n = 4096
    items = torch.arange(1, n + 1, dtype=torch.int64, device="cuda") * 3
    q.enqueue(items, serial=True)
    out, got = q.dequeue(n, serial=True)
    torch.cuda.synchronize()
    assert bool(got.all())
    assert torch.equal(out, items)
    q.destroy()
    _done(a)


def test_queue_concurrent_producers_and_consumers(gpu):
    import torch
    a = _alloc(gpu)
    q = gpu.Queue(a)
    n = 1 << 19
    items = torch.arange(1, n + 1, dtype=torch.int64, device="cuda")
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    with torch.cuda.stream(s1):
        ok = q.enqueue(items)
    with torch.cuda.stream(s2):
        out1, got1 = q.dequeue(n)
    torch.cuda.synchronize()
    assert bool(ok.all())
    first = out1[got1.bool()]
    out2, got2 = q.dequeue(n)
    torch.cuda.synchronize()
    second = out2[got2.bool()]
    allv = np.sort(torch.cat([first, second]).cpu().numpy())
    assert np.array_equal(allv, np.arange(1, n + 1)), "items lost or duplicated"
    q.destroy()
    _done(a)


def test_queue_ring(gpu):
    import torch
    a = _alloc(gpu)
    q = gpu.Queue(a, ring=True)
    n = 1 << 16
    items = torch.arange(1, n + 1, dtype=torch.int64, device="cuda")
    assert bool(q.enqueue(items).all())
    taken = []
    for _ in range(64):
        out, got = q.dequeue(n)
        torch.cuda.synchronize()
        taken.append(out[got.bool()].cpu().numpy())
        if sum(len(t) for t in taken) == n:
            break
    allv = np.sort(np.concatenate(taken))
    assert np.array_equal(allv, np.arange(1, n + 1))
    q.destroy()
    _done(a)

"""Block queue on the global allocator (block_queue.cuh:240-463).

The reference's test (tests/src/block_queue_test.cu:27-33) enqueues every
thread id in parallel; its dequeue is commented out, so dequeue is pinned
only by invariants here: every item enqueued comes out exactly once, with
producers and consumers running at the same time too; one producer's items
come out in its order; an empty queue reports empty. Nodes are never
reused while the queue runs, so n enqueues fill exactly ceil(n / size)
nodes; the nodes that dequeues empty go to the retired list, reclaim
returns them, and after destroy the allocator drains to its fresh state.
"""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def glob(gpu):
    gpu.init_global_allocator(8 << 30, 31, print_info=False)
    yield gpu
    g = gpu.global_allocator()
    g.drain()
    H.check_fresh(g)
    gpu.free_global_allocator()


@pytest.mark.parametrize("kind", [0, 1])
def test_bq_parallel_enqueue_dequeue(glob, kind):
    import torch
    n = 1 << 20 if kind == 0 else 1 << 18  # 16-item nodes: every node change is contended
    q = glob.BlockQueue(kind)
    size = q.items_per_block
    ok = q.enqueue(torch.arange(n, dtype=torch.int64, device="cuda"))
    torch.cuda.synchronize()
    assert bool(ok.all())
    assert q.stats() == {"nodes": -(-n // size), "retired": 0, "failed": 0}
    out, ok = q.dequeue(n)
    torch.cuda.synchronize()
    assert bool(ok.all())
    assert np.array_equal(np.sort(out.cpu().numpy()), np.arange(n))
    _, ok = q.dequeue(1 << 12)
    torch.cuda.synchronize()
    assert not bool(ok.any())
    st = q.stats()
    assert st["nodes"] + st["retired"] == -(-n // size) and st["nodes"] >= 1
    assert q.reclaim() == st["retired"]
    assert q.stats() == {"nodes": st["nodes"], "retired": 0, "failed": 0}
    q.destroy()


@pytest.mark.parametrize("kind", [0, 1])
def test_bq_single_producer_order(glob, kind):
    import torch
    n = 5000
    items = torch.arange(7, 7 + n, dtype=torch.int64, device="cuda")
    q = glob.BlockQueue(kind)
    ok = q.enqueue(items, serial=True)
    out, ok2 = q.dequeue(n + 10, serial=True)
    torch.cuda.synchronize()
    assert bool(ok.all())
    o, k = out.cpu().numpy(), ok2.cpu().numpy()
    assert k[:n].all() and not k[n:].any()
    assert np.array_equal(o[:n], np.arange(7, 7 + n))
    q.destroy()


@pytest.mark.parametrize("kind", [0, 1])
def test_bq_producers_and_consumers_together(glob, kind):
    """Enqueues on one stream while dequeues run on another; what the
    consumers missed is dequeued afterwards: every item exactly once."""
    import torch
    n = 1 << 19 if kind == 0 else 1 << 17
    q = glob.BlockQueue(kind)
    s_p, s_c = torch.cuda.Stream(), torch.cuda.Stream()
    items = torch.arange(1, n + 1, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    with torch.cuda.stream(s_p):
        ok = q.enqueue(items)
    outs = []
    with torch.cuda.stream(s_c):
        for _ in range(4):
            outs.append(q.dequeue(n // 4))
    torch.cuda.synchronize()
    assert bool(ok.all())
    got = [o[k.bool()] for o, k in outs]
    rest, ok_rest = q.dequeue(n)
    torch.cuda.synchronize()
    got.append(rest[ok_rest.bool()])
    allv = torch.cat(got).cpu().numpy()
    assert len(allv) == n
    assert np.array_equal(np.sort(allv), np.arange(1, n + 1))
    q.destroy()


def test_bq_rounds_with_reclaim(glob):
    """Rounds of enqueue and dequeue with reclaim between them keep only
    the nodes of the items still queued, and destroy with items left
    returns every node."""
    import torch
    q = glob.BlockQueue(1)
    base = 0
    for r in range(6):
        n = 10000 + 1000 * r
        ok = q.enqueue(torch.arange(base, base + n, dtype=torch.int64, device="cuda"))
        out, ok2 = q.dequeue(n - 100)
        torch.cuda.synchronize()
        assert bool(ok.all()) and bool(ok2.all())
        q.reclaim()
        st = q.stats()
        left = 100 * (r + 1)
        assert st["retired"] == 0 and st["nodes"] <= -(-left // 16) + 1
        base += n
    q.destroy()

"""Serving loop on the global allocator: sessions grow allocator-backed
chunked vectors (128 B doubling to 64 KiB, then 64 KiB chunks) and are
freed when they end, a hash table grows beside them; every word of every
live session reads back, the counters agree with each other, no append is
cut short while memory is there, admission stops when the free-segment
reserve is reached, chunk allocations that cannot be served fail at once
(counted, the step still ends), and after the loop is torn down the
allocator drains to its fresh state (nothing leaked).  The serving loop has
no reference counterpart: it is pinned by these invariants only."""
import time

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu


def _check_counts(r, steps, ht_ops):
    c = r["counters"]
    assert r["session_words_bad"] == 0
    assert r["live_sessions"] == c["admitted"] - c["ended"]
    assert c["ht_inserted"] + c["ht_removed"] + c["ht_absent"] + c["ht_failed"] == steps * ht_ops
    assert c["words"] > 0 and c["chunks"] > 0


def test_serving_small(gpu):
    import torch
    from gallatin_amd import serving as SV
    gpu.init_global_allocator(16 << 30, 9, print_info=False)
    try:
        r = SV.run(torch, steps=40, warmup=20, seed=3, max_sessions=1 << 15,
                   arrivals_mean=300, p_end=0.05, ht_ops=1 << 14, reserve_fraction=0.0)
        c = r["counters"]
        _check_counts(r, 60, 1 << 14)
        assert c["chunk_failed"] == 0 and c["ht_failed"] == 0 and c["refused"] == 0
        assert c["words_refused"] == 0 and c["chunks_freed"] > 0
        # the words read back are exactly those appended and not yet freed
        assert r["session_words"] <= c["words"]
        g = gpu.global_allocator()
        g.drain()
        H.check_fresh(g)
    finally:
        gpu.free_global_allocator()


def test_serving_sessions_grow_past_the_doubling_chunks(gpu):
    """Long-lived sessions hold more than the 128 KiB of the doubling
    chunks: later 64 KiB chunks and their index are allocated, written,
    read back and freed."""
    import torch
    from gallatin_amd import serving as SV
    gpu.init_global_allocator(16 << 30, 21, print_info=False)
    try:
        r = SV.run(torch, steps=30, warmup=10, seed=5, max_sessions=1 << 12,
                   arrivals_mean=100, p_end=0.002, ht_ops=0, reserve_fraction=0.0)
        c = r["counters"]
        _check_counts(r, 40, 0)
        assert c["words_refused"] == 0 and c["chunk_failed"] == 0
        # sessions of 10+ steps hold about 10,000+ words each; the first ones
        # (40 steps) are past 16384 words
        assert r["session_words"] > 16384 * 100
        assert r["gb_live"] * 1e9 >= r["session_words"] * 8
        g = gpu.global_allocator()
        g.drain()
        H.check_fresh(g)
    finally:
        gpu.free_global_allocator()


def test_serving_admission_gate(gpu):
    """A reserve above the pool's segment count refuses every arrival."""
    import torch
    from gallatin_amd import serving as SV
    gpu.init_global_allocator(4 << 30, 9, print_info=False)
    try:
        r = SV.run(torch, steps=5, warmup=0, seed=4, max_sessions=1 << 12,
                   arrivals_mean=100, p_end=0.05, ht_ops=0, reserve_fraction=2.0)
        assert r["counters"]["admitted"] == 0 and r["counters"]["refused"] > 0
        g = gpu.global_allocator()
        g.drain()
        H.check_fresh(g)
    finally:
        gpu.free_global_allocator()


def test_serving_out_of_memory_fails_fast(gpu):
    """On a 4 GiB pool with no admission reserve, 8192 long-lived sessions
    outgrow the pool (about 68 MB of appends per step): chunk allocations
    fail, the appends stop before the missing chunk (words refused), every
    word written still reads back, the steps end promptly, and teardown
    drains to fresh."""
    import torch
    from gallatin_amd import serving as SV
    gpu.init_global_allocator(4 << 30, 13, print_info=False)
    try:
        t0 = time.perf_counter()
        r = SV.run(torch, steps=60, warmup=10, seed=6, max_sessions=1 << 13,
                   arrivals_mean=800, p_end=0.002, ht_ops=1 << 10, reserve_fraction=0.0)
        dt = time.perf_counter() - t0
        c = r["counters"]
        _check_counts(r, 70, 1 << 10)
        assert c["chunk_failed"] > 0 and c["words_refused"] > 0
        assert dt < 30, dt
        g = gpu.global_allocator()
        assert g.info()["errors"] == 0
        g.drain()
        H.check_fresh(g)
    finally:
        gpu.free_global_allocator()


@pytest.mark.parametrize("seed", [1234, 1238, 1239, 1240])
def test_serving_full_config(gpu, seed):
    """The serving configuration at full size: 200 GiB pool, Poisson(2000)
    arrivals, geometric(0.01) lifetimes, 2^20 table operations per step on
    keys in [1, 2^28], a 10 % free-segment reserve.  Over 700 steps the pool
    fills, the gate closes (arrivals refused) and opens again; no append is
    cut short, every word reads back, and teardown drains to fresh.  Several
    seeds: block claims that livelocked on a segment's last blocks (dozens
    of waves subtracting and adding back on one count) failed chunk
    allocations in about a third of the runs, at any seed."""
    import torch
    from gallatin_amd import serving as SV
    gpu.init_global_allocator(200 << 30, seed, print_info=False)
    try:
        r = SV.run(torch, steps=50, warmup=650, seed=seed, max_sessions=1 << 19,
                   arrivals_mean=2000, p_end=0.01, ht_ops=1 << 20, reserve_fraction=0.10,
                   trace_every=50)
        c = r["counters"]
        _check_counts(r, 700, 1 << 20)
        assert c["words_refused"] == 0 and c["chunk_failed"] == 0 and c["ht_failed"] == 0, \
            (c, r["allocator_events"], r["trace"])
        assert c["refused"] > 0, "the admission gate never closed"
        assert min(t["segments_free"] for t in r["trace"]) >= r["reserve_segments"] // 2
        g = gpu.global_allocator()
        assert g.info()["errors"] == 0
        g.drain()
        H.check_fresh(g)
    finally:
        gpu.free_global_allocator()


def test_vector_chunk_failure_returns_at_once(gpu):
    """fixed_vector on a pool with no segment left for its 16 MiB chunk
    (positions [2^21, 2^22)): inserts into it return ~0 at once instead of
    waiting out the bounded poll; a later insert into the same chunk tries
    the allocation again (the failed claim was given up) and fails fast
    too; once memory is freed, the chunk is allocated and used."""
    import torch
    gpu.init_global_allocator(1 << 30, 7, print_info=False)
    try:
        # take every free segment with whole 16 MiB blocks
        sizes = torch.full((64,), 16 << 20, dtype=torch.int64, device="cuda")
        held = gpu.global_malloc_bulk(sizes)
        torch.cuda.synchronize()
        assert int((held.cpu() == 0).sum()) > 0  # the pool ran out
        v = gpu.Vector()
        fail = np.uint64(0xFFFFFFFFFFFFFFFF)

        def insert(n, first):
            items = torch.arange(first, first + n, dtype=torch.int64, device="cuda")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pos = v.insert(items, bulk=True)
            torch.cuda.synchronize()
            return pos.cpu().numpy().view(np.uint64), time.perf_counter() - t0

        head = 1 << 21
        p, dt0 = insert(head + 1000, 1)
        assert int((p == fail).sum()) == 1000
        assert np.array_equal(np.sort(p[p != fail]), np.arange(head, dtype=np.uint64))
        assert dt0 < 5.0, dt0
        # a claim left set would hold each of these 16 waves for the whole
        # poll (2^22 loads and sleeps, seconds); a released one lets them fail
        # after their own allocation attempt
        p, dt1 = insert(1000, head + 1001)
        assert np.all(p == fail)
        assert dt1 < 1.0, (dt0, dt1)
        gpu.global_free_bulk(held)
        torch.cuda.synchronize()
        p, dt = insert(1000, head + 2001)
        assert np.all(p != fail)
        assert np.array_equal(np.sort(p), np.arange(head + 2000, head + 3000, dtype=np.uint64))
        v.destroy()
        g = gpu.global_allocator()
        assert g.info()["errors"] == 0
        g.drain()
        H.check_fresh(g)
    finally:
        gpu.free_global_allocator()

"""Serving loop on the global allocator: sessions grow allocator-backed
vectors and are freed when they end, a hash table grows beside them; every
word of every live session reads back, the counters agree with each other,
admission stops when the free-segment reserve is reached, and after the
loop is torn down the allocator drains to its fresh state (nothing leaked)."""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu


def test_serving_small(gpu):
    import torch
    from gallatin_amd import serving as SV
    gpu.init_global_allocator(16 << 30, 9, print_info=False)
    try:
        r = SV.run(torch, steps=40, warmup=20, seed=3, max_sessions=1 << 15,
                   arrivals_mean=300, p_end=0.05, ht_ops=1 << 14, reserve_fraction=0.0)
        c = r["counters"]
        assert r["session_words_bad"] == 0
        assert r["live_sessions"] == c["admitted"] - c["ended"]
        assert c["chunk_failed"] == 0 and c["ht_failed"] == 0 and c["refused"] == 0
        assert c["ht_inserted"] + c["ht_removed"] + c["ht_absent"] == 60 * (1 << 14)
        assert c["words"] > 0 and c["chunks"] > 0 and c["chunks_freed"] > 0
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

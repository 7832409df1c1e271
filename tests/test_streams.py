"""The block and multi-segment request stream (include/gallatin/streams.h,
gal_big_*) on the CPU, through the oracle library, which compiles the same
header as the GPU kernels:

- the free order of every step is a bijection of [0, n): each request is
  freed exactly once, for small, ragged and power-of-two n;
- sizes lie in [64 KiB, 64 MiB] for the first n_small requests and in
  [256 MiB, 2 GiB] for the rest, spread over the whole range;
- the glibc baseline, which runs that stream on host threads, frees every
  request it served and finds every tag intact;
- the glibc baselines of the plain small stream and of a batched Zipf
  stream do the same.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KiB, MiB, GiB = 1 << 10, 1 << 20, 1 << 30


@pytest.mark.parametrize("n_small,n_large", [(1, 0), (0, 1), (2, 0), (3, 0), (63, 1),
                                             (64, 0), (64, 1), (1000, 0), (4096, 64),
                                             (8192, 0), (8191, 2)])
def test_free_order_is_a_bijection(n_small, n_large):
    n = n_small + n_large
    orders = []
    for step in range(4):
        _, order = oracle.big_stream(7, step, n_small, n_large)
        assert np.array_equal(np.sort(order), np.arange(n, dtype=np.uint64))
        orders.append(order)
    if n >= 64:
        # a shuffled order: few fixed points, and each step's order differs
        assert np.sum(orders[0] == np.arange(n)) < n // 8
        assert not np.array_equal(orders[0], orders[1])


def test_sizes_cover_their_ranges():
    sizes, _ = oracle.big_stream(11, 0, 4096, 64)
    small, large = sizes[:4096].astype(np.float64), sizes[4096:].astype(np.float64)
    assert small.min() >= 64 * KiB and small.max() <= 64 * MiB
    assert large.min() >= 256 * MiB and large.max() <= 2 * GiB
    # uniform: the mean of 4096 draws is within 2 MiB of the midpoint
    # (standard error 0.29 MiB), and every eighth of the range is used
    assert abs(small.mean() - (64 * KiB + 64 * MiB) / 2) < 2 * MiB
    edges = np.linspace(64 * KiB, 64 * MiB, 9)
    assert np.all(np.histogram(small, edges)[0] > 300)
    s1, _ = oracle.big_stream(11, 1, 4096, 64)
    assert not np.array_equal(sizes, s1)


def test_glibc_baseline_frees_every_request():
    from gallatin_amd import cpu_baseline as C
    r = C.big(3, 0, 3, 509, 3, 4)
    assert r["misses"] == 0 and r["bad"] == 0
    assert r["pairs"] == 3 * 512 and r["bytes"] == 2 * 8 * 3 * 512 and r["threads"] == 4


def test_glibc_plain_baseline():
    from gallatin_amd import cpu_baseline as C
    r = C.plain(10000, 3, 4)
    assert r["pairs"] == 30000 and r["misses"] == 0 and r["bad"] == 0 and r["threads"] == 4


def test_glibc_batches_baseline():
    """Every malloc of a small Zipf stream is freed exactly once (in a later
    batch or at the end) with its tag intact."""
    from gallatin_amd import batch_streams as B
    from gallatin_amd import cpu_baseline as C
    ops, off, nslots = B.zipf_stream(11, nbatches=48, max_count=1 << 13)
    mallocs, _ = B.check_stream(ops, off, nslots)
    r = C.batches(ops, off, nslots, 4)
    assert r["misses"] == 0 and r["bad"] == 0
    assert r["pairs"] == mallocs


def test_glibc_serving_baseline():
    """A small serving run on host threads: sessions live, append and end,
    the table holds keys, and every ending session's first word is intact.
    Sessions grow past 128 KiB (the doubling chunks) into 64 KiB chunks."""
    from gallatin_amd import cpu_baseline as C
    from gallatin_amd import serving as S
    r, ms = C.serving(5, 10, 3, 4096, 200.0, 0.05, 4096, S.tokens_table(), 1 << 40, 0.05, 4)
    assert r["bad"] == 0 and r["live"] > 0 and r["ended"] > 0 and r["table_keys"] > 0
    assert r["appended_words"] > 0 and r["table_ops"] == 13 * 4096 and np.all(ms > 0)
    assert r["refused"] == 0 and r["admitted"] > 0
    # 512 slots filled in about 10 steps, then 20 more steps of about 1056
    # words each: most sessions are past the 16384 words of the doubling
    # chunks and hold 64 KiB chunks beyond them
    r, _ = C.serving(7, 0, 30, 512, 50.0, 0.01, 0, S.tokens_table(), 1 << 40, 0.05, 4)
    assert r["bad"] == 0 and r["live"] > 400
    assert r["held_bytes"] > r["live"] * 16384 * 8 // 2


def test_glibc_serving_gate():
    """The host gate refuses arrivals once the chunk bytes held reach the
    pool budget less its reserve, and admits none when the budget is 0."""
    from gallatin_amd import cpu_baseline as C
    from gallatin_amd import serving as S
    r, _ = C.serving(5, 0, 8, 4096, 200.0, 0.05, 0, S.tokens_table(), 0, 0.05, 4)
    assert r["admitted"] == 0 and r["refused"] > 0 and r["live"] == 0
    # a budget of about 20 sessions' first steps: admission stops early
    r, _ = C.serving(5, 2, 10, 4096, 200.0, 0.01, 0, S.tokens_table(), 24 << 20, 0.05, 4)
    assert r["refused"] > 0 and r["bad"] == 0
    assert r["held_bytes"] < (24 << 20) * 4

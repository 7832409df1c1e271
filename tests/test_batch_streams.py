"""Host generator of batched request streams (CPU)."""
import numpy as np
import pytest

from gallatin_amd import batch_streams as B
from gallatin_amd.batch_streams import zipf_stream


def test_bounded_zipf_range_and_head():
    rng = np.random.default_rng(0)
    x = B.bounded_zipf(rng, 200000, 1.2, 1 << 20)
    assert x.min() >= 1 and x.max() <= 1 << 20
    # P(X = 1) = 1 / H(2^20, 1.2) for the bounded law
    k = np.arange(1, (1 << 20) + 1, dtype=np.float64)
    p1 = 1.0 / np.sum(k ** -1.2)
    assert abs((x == 1).mean() - p1) < 0.01


def test_lognormal_sizes_clipped():
    rng = np.random.default_rng(1)
    s = B.lognormal_sizes(rng, 100000, 6.0, 2.0, 16 << 20)
    assert s.min() >= 1 and s.max() <= 16 << 20
    assert abs(np.median(s) - np.exp(6.0)) < 25   # median e^mu = 403


def test_zipf_stream_structure():
    ops, off, nslots = B.zipf_stream(3, nbatches=64, max_count=1 << 14)
    mallocs, frees = B.check_stream(ops, off, nslots)
    assert mallocs + frees == len(ops) == int(off[-1])
    assert frees > 0
    frac = mallocs / len(ops)
    assert 0.55 < frac < 0.75    # 60 % mallocs, plus frees with nothing live


def test_zipf_stream_reproducible():
    a = B.zipf_stream(5, nbatches=16, max_count=4096)
    b = B.zipf_stream(5, nbatches=16, max_count=4096)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_check_stream_catches_a_double_free():
    ops, off, nslots = B.zipf_stream(4, nbatches=32, max_count=4096)
    bad = ops.copy()
    frees = np.nonzero(bad["kind"] == 1)[0]
    assert len(frees) >= 2
    bad["slot"][frees[-1]] = bad["slot"][frees[0]]   # free the same slot twice
    with pytest.raises(AssertionError):
        B.check_stream(bad, off, nslots)


def _ops(rows):
    import gallatin_amd as G
    a = np.zeros(len(rows), dtype=G.BATCH_OP_DTYPE)
    for i, (kind, slot) in enumerate(rows):
        a[i] = (16, kind, slot)
    return a


def test_stream_check_accepts_the_zipf_stream():
    import gallatin_amd as G
    ops, off, nslots = zipf_stream(5, nbatches=32, max_count=1 << 10)
    G.check_batch_stream(ops, off, nslots, persistent=True)


@pytest.mark.parametrize("rows,offsets,persistent,msg", [
    ([(0, 0), (0, 0)], [0, 1, 2], True, "filled a second time"),
    ([(0, 0), (1, 0), (0, 0)], [0, 1, 2, 3], True, "filled a second time"),
    ([(0, 0), (0, 0)], [0, 1, 2], False, "before its pointer was freed"),
    ([(0, 0), (1, 0)], [0, 2], False, "without a malloc in an earlier batch"),
    ([(1, 3)], [0, 1], False, "without a malloc in an earlier batch"),
    ([(0, 9)], [0, 1], False, "slot beyond nslots"),
])
def test_stream_check_refuses(rows, offsets, persistent, msg):
    """The plan's slot rules are checked on the host: a slot filled twice in
    a persistent stream, filled again before its free, freed in its own
    batch or never filled, or out of range is refused with its op named."""
    import gallatin_amd as G
    with pytest.raises(G.GallatinError, match=msg):
        G.check_batch_stream(_ops(rows), offsets, 4, persistent=persistent)


def test_stream_check_allows_reuse_across_batches_eagerly():
    """Per-batch launches may refill a slot after its free in a later batch."""
    import gallatin_amd as G
    G.check_batch_stream(_ops([(0, 0), (1, 0), (0, 0), (1, 0)]), [0, 1, 2, 3, 4], 4,
                         persistent=False)

"""Allocator-backed data structures on the GPU (global allocator).

fixed_vector: every insert gets a distinct position and the item lands
there (the uniqueness check of the reference's tests/src/vector_tests.cu:
121-137); bulk inserts likewise; inserts beyond max_items return ~0.

extendible hash table: query-after-insert (tests/src/extendible_ht_test.cu:
223-249) through many bucket splits; replace, remove and re-insert; keys
that were never inserted are absent; queries of present keys running
while another stream inserts and splits buckets all succeed.  remove has
no reference counterpart, so its results are pinned only by these
invariants.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FAIL = np.uint64(0xFFFFFFFFFFFFFFFF)


@pytest.fixture(scope="module")
def glob(gpu):
    gpu.init_global_allocator(8 << 30, 17, print_info=False)
    yield gpu
    gpu.free_global_allocator()


def _keys(rng, n):
    """n distinct keys, none of them 0 or ~0."""
    k = np.unique(rng.integers(1, 1 << 62, int(n * 1.1), dtype=np.int64))[:n]
    assert len(k) == n
    return rng.permutation(k)


@pytest.mark.parametrize("bulk", [False, True])
def test_vector_positions_unique(glob, bulk):
    import torch
    n = 1 << 20
    items = torch.arange(1, n + 1, dtype=torch.int64, device="cuda")
    v = glob.Vector()
    pos = v.insert(items, bulk=bulk)
    torch.cuda.synchronize()
    p = pos.cpu().numpy().view(np.uint64)
    assert np.all(p != FAIL)
    assert len(np.unique(p)) == n and p.max() == n - 1
    out = v.export()
    assert len(out) == n
    assert np.array_equal(out[p.astype(np.int64)], np.arange(1, n + 1, dtype=np.uint64))
    v.destroy()


def test_vector_full(glob):
    import torch
    cap, extra = 1 << 22, 1000
    items = torch.arange(1, cap + extra + 1, dtype=torch.int64, device="cuda")
    v = glob.Vector()
    pos = v.insert(items, bulk=True)
    torch.cuda.synchronize()
    p = pos.cpu().numpy().view(np.uint64)
    assert int((p == FAIL).sum()) == extra
    ok = p[p != FAIL]
    assert len(np.unique(ok)) == cap
    out = v.export()
    assert len(out) == cap and len(np.unique(out)) == cap
    v.destroy()


def test_table_insert_query_remove(glob):
    import torch
    rng = np.random.default_rng(3)
    n = 1 << 17
    keys = _keys(rng, n)
    vals = rng.integers(0, 1 << 62, n, dtype=np.int64)
    t = glob.Table(kind=0)
    dk, dv = torch.from_numpy(keys).cuda(), torch.from_numpy(vals).cuda()
    ok = t.insert(dk, dv)
    torch.cuda.synchronize()
    assert bool(ok.all())
    got, found = t.query(dk)
    torch.cuda.synchronize()
    assert bool(found.all())
    assert np.array_equal(got.cpu().numpy(), vals)
    st = t.stats()
    assert st["keys"] == n and st["failed"] == 0
    assert st["level"] > 4 and st["buckets"] > 16      # the table split many times
    # replace every value
    t.insert(dk, dv + 1)
    got, found = t.query(dk)
    torch.cuda.synchronize()
    assert bool(found.all()) and np.array_equal(got.cpu().numpy(), vals + 1)
    # remove 30 %
    gone = rng.random(n) < 0.3
    rm = t.remove(torch.from_numpy(keys[gone]).cuda())
    torch.cuda.synchronize()
    assert bool(rm.all())
    got, found = t.query(dk)
    torch.cuda.synchronize()
    f = found.cpu().numpy().astype(bool)
    assert not np.any(f[gone]) and np.all(f[~gone])
    assert np.array_equal(got.cpu().numpy()[~gone], vals[~gone] + 1)
    assert t.stats()["keys"] == int((~gone).sum())
    rm2 = t.remove(torch.from_numpy(keys[gone]).cuda())   # already gone
    torch.cuda.synchronize()
    assert not bool(rm2.any())
    # absent keys
    other = _keys(np.random.default_rng(99), 4096) | np.int64(1 << 62)
    _, found = t.query(torch.from_numpy(other).cuda())
    torch.cuda.synchronize()
    assert not bool(found.any())
    # re-insert the removed keys
    t.insert(torch.from_numpy(keys[gone]).cuda(), torch.from_numpy(vals[gone] + 7).cuda())
    got, found = t.query(dk)
    torch.cuda.synchronize()
    assert bool(found.all())
    exp = np.where(gone, vals + 7, vals + 1)
    assert np.array_equal(got.cpu().numpy(), exp)
    t.destroy()


def test_table_queries_during_splits(glob):
    """Present keys are queried on one stream while another stream inserts
    four times as many new keys (forcing splits of the buckets that hold
    them): every query finds its key with its value."""
    import torch
    rng = np.random.default_rng(4)
    n0, n1 = 1 << 15, 1 << 17
    keys = _keys(rng, n0 + n1)
    k0, k1 = keys[:n0], keys[n0:]
    t = glob.Table(kind=0)
    d0 = torch.from_numpy(k0).cuda()
    t.insert(d0, d0 * 3)
    torch.cuda.synchronize()
    s_ins, s_q = torch.cuda.Stream(), torch.cuda.Stream()
    d1 = torch.from_numpy(k1).cuda()
    reps = []
    with torch.cuda.stream(s_ins):
        ok = t.insert(d1, d1 * 5)
    with torch.cuda.stream(s_q):
        for _ in range(8):
            reps.append(t.query(d0))
    torch.cuda.synchronize()
    assert bool(ok.all())
    for got, found in reps:
        assert bool(found.all())
        assert torch.equal(got, d0 * 3)
    st = t.stats()
    assert st["keys"] == n0 + n1 and st["failed"] == 0
    t.destroy()


def test_large_table(glob):
    import torch
    rng = np.random.default_rng(5)
    n = 1 << 20
    keys = _keys(rng, n)
    t = glob.Table(kind=1)
    dk = torch.from_numpy(keys).cuda()
    ok = t.insert(dk, dk ^ 0x5555)
    got, found = t.query(dk)
    torch.cuda.synchronize()
    assert bool(ok.all()) and bool(found.all())
    assert torch.equal(got, dk ^ 0x5555)
    assert t.stats()["keys"] == n
    t.destroy()

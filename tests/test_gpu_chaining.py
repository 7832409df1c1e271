"""Chained hash table on the global allocator (chaining_table.cuh:114-335).

The reference's test (tests/src/chaining_test.cu:38-79) inserts keys tid+1
with values tid and queries every key back.  Beyond that, the chain of each
key is fixed by the reference's hash (MurmurHash64A over the key's bytes,
modulo the number of chains, chaining_table.cuh:244), and since slots are
never emptied a chain that received c keys holds exactly ceil(c / size)
blocks: the per-chain block counts are compared with that count, taken
from the oracle's hash.  Duplicates are stored and the first one wins;
queries of present keys running while another stream inserts all succeed
with their values; after the table is destroyed the allocator drains to its
fresh state.
"""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

KINDS = {0: (np.uint64, 4), 1: (np.uint64, 16), 2: (np.uint32, 4), 3: (np.uint32, 32),
         4: (np.uint64, 1)}
SEED = 123456


@pytest.fixture(scope="module")
def glob(gpu):
    gpu.init_global_allocator(8 << 30, 23, print_info=False)
    yield gpu
    g = gpu.global_allocator()
    g.drain()
    H.check_fresh(g)
    gpu.free_global_allocator()


def _torch(a):
    import torch
    t = torch.from_numpy(a.view(np.int64 if a.dtype == np.uint64 else np.int32).copy())
    return t.cuda()


def _expected_lengths(oracle_mod, keys, nblocks, size, prealloc):
    chains = oracle_mod.chain_of(keys, SEED, nblocks)
    counts = np.bincount(chains.astype(np.int64), minlength=nblocks)
    exp = (counts + size - 1) // size
    return np.maximum(exp, 1) if prealloc else exp


@pytest.mark.parametrize("prealloc", [False, True])
@pytest.mark.parametrize("kind", sorted(KINDS))
def test_chain_insert_query(glob, oracle_mod, kind, prealloc):
    import torch
    dt, size = KINDS[kind]
    n = 1 << 17
    keys = np.arange(1, n + 1, dtype=dt)
    vals = np.arange(n, dtype=dt)
    t = glob.ChainTable(kind, nslots=n, default_key=0, seed=SEED, prealloc=prealloc)
    ok = t.insert(_torch(keys), _torch(vals))
    got, found = t.query(_torch(keys))
    torch.cuda.synchronize()
    assert bool(ok.all())
    assert bool(found.all())
    assert np.array_equal(got.cpu().numpy().view(dt), vals)
    # never-inserted keys and the empty-slot key are absent; the empty-slot
    # key cannot be inserted
    _, found = t.query(_torch(np.arange(n + 1, 2 * n + 1, dtype=dt)))
    ok0 = t.insert(_torch(np.zeros(4, dtype=dt)), _torch(np.ones(4, dtype=dt)))
    _, found0 = t.query(_torch(np.zeros(4, dtype=dt)))
    torch.cuda.synchronize()
    assert not bool(found.any()) and not bool(ok0.any()) and not bool(found0.any())
    nblocks = n // size
    lengths = t.lengths().cpu().numpy()
    exp = _expected_lengths(oracle_mod, keys, nblocks, size, prealloc)
    assert np.array_equal(lengths, exp)
    st = t.stats()
    assert st == {"max_chain": int(exp.max()), "blocks": int(exp.sum()), "nblocks": nblocks,
                  "failed": 0}
    t.destroy()


def test_chain_duplicates_first_wins(glob, oracle_mod):
    """A key inserted twice is stored twice; query returns the first."""
    import torch
    dt, size = KINDS[0]
    keys = np.arange(1, 4097, dtype=dt)
    t = glob.ChainTable(0, nslots=1024, seed=SEED)
    t.insert(_torch(keys), _torch(keys * 10))
    torch.cuda.synchronize()
    t.insert(_torch(keys), _torch(keys * 100))
    got, found = t.query(_torch(keys))
    torch.cuda.synchronize()
    assert bool(found.all())
    assert np.array_equal(got.cpu().numpy().view(dt), keys * 10)
    exp = _expected_lengths(oracle_mod, np.concatenate([keys, keys]), 1024 // size, size, False)
    assert np.array_equal(t.lengths().cpu().numpy(), exp)
    t.destroy()


@pytest.mark.parametrize("kind", [0, 3])
def test_chain_queries_during_inserts(glob, kind):
    """Present keys are queried on one stream while another stream inserts
    four times as many keys into the same chains: every query finds its key
    with its value."""
    import torch
    dt, size = KINDS[kind]
    n0, n1 = 1 << 15, 1 << 17
    keys = np.arange(1, n0 + n1 + 1, dtype=dt)
    k0, k1 = keys[:n0], keys[n0:]
    t = glob.ChainTable(kind, nslots=n0, seed=SEED)
    d0 = _torch(k0)
    t.insert(d0, _torch(k0 * 3))
    torch.cuda.synchronize()
    s_ins, s_q = torch.cuda.Stream(), torch.cuda.Stream()
    d1, v1 = _torch(k1), _torch(k1 * 5)
    reps = []
    with torch.cuda.stream(s_ins):
        ok = t.insert(d1, v1)
    with torch.cuda.stream(s_q):
        for _ in range(8):
            reps.append(t.query(d0))
    torch.cuda.synchronize()
    assert bool(ok.all())
    for got, found in reps:
        assert bool(found.all())
        assert np.array_equal(got.cpu().numpy().view(dt), k0 * 3)
    got, found = t.query(d1)
    torch.cuda.synchronize()
    assert bool(found.all()) and np.array_equal(got.cpu().numpy().view(dt), k1 * 5)
    t.destroy()


def test_chain_refuses_tiny_table(glob):
    with pytest.raises(glob.GallatinError):
        glob.ChainTable(1, nslots=8)

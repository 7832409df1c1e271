"""Pins the oracle (CPU): published hash KAT, closed forms, brute force on
small trees, hand-derived known answers and exact allocator invariants."""
import struct

import numpy as np
import pytest

import oracle
from tests import helpers as H

G = H.golden()


# ------------------------------------------------------------ hash / bits

def test_murmur64a_smhasher_verification():
    hashes = b"".join(struct.pack("<Q", oracle.murmur64a(bytes(range(i)), 256 - i))
                      for i in range(256))
    final = oracle.murmur64a(hashes, 0) & 0xFFFFFFFF
    assert final == int(G["murmur64a_smhasher_verification"], 16)


@pytest.mark.parametrize("dt,fmt", [(np.uint32, "<I"), (np.uint64, "<Q")])
def test_chain_of_is_murmur_of_key_bytes(dt, fmt):
    """chain_of (the chaining table's bucket choice) hashes each key's
    little-endian bytes with the KAT-pinned murmur64a, modulo the chains."""
    rng = np.random.default_rng(5)
    keys = rng.integers(0, np.iinfo(dt).max, 500, dtype=dt)
    got = oracle.chain_of(keys, 123456, 997)
    exp = [oracle.murmur64a(struct.pack(fmt, int(k)), 123456) % 997 for k in keys]
    assert np.array_equal(got, np.array(exp, dtype=np.uint64))


def test_clog2_closed_form():
    for x in list(range(1, 5000)) + [2 ** k + d for k in range(12, 63) for d in (-1, 0, 1)]:
        assert oracle.clog2(x) == H.clog2(x), x


def test_cfcll_brute_force():
    rng = np.random.default_rng(1)
    vals = [0, 1, 2, 3, (1 << 64) - 1, (1 << 63), (1 << 63) - 1]
    vals += [int(v) for v in rng.integers(0, 2 ** 63, 200, dtype=np.int64)]
    vals += [(1 << k) - 1 for k in range(65)]
    for v in vals:
        v &= (1 << 64) - 1
        n = 0
        while n < 64 and (v >> n) & 1:
            n += 1
        assert oracle.cfcll(v) == n, hex(v)


# ------------------------------------------------------------ vEB trees

def _layers_consistent(t):
    layers = t.layers()
    for li in range(len(layers) - 1):
        parent, child = layers[li], layers[li + 1]
        for w, word in enumerate(child):
            bit = (int(parent[w // 64]) >> (w % 64)) & 1
            assert bit == (int(word) != 0), (li, w)


@pytest.mark.parametrize("universe", [1, 5, 63, 64, 65, 130, 4096, 4097, 12800])
def test_veb_fresh_layout(universe):
    t = oracle.Veb(universe, filled=True)
    exp = H.fresh_layers(universe, True)
    got = t.layers()
    assert len(got) == len(exp)
    for a, b in zip(got, exp):
        assert np.array_equal(a, b)
    _layers_consistent(t)


@pytest.mark.parametrize("universe", [1, 7, 64, 65, 200, 4097, 9000])
def test_veb_random_ops_vs_set(universe):
    rng = np.random.default_rng(universe)
    t = oracle.Veb(universe, filled=False)
    ref = set()
    for _ in range(3000):
        op = rng.integers(0, 6)
        x = int(rng.integers(0, universe))
        if op <= 1:
            assert t.insert(x) == (x not in ref)
            ref.add(x)
        elif op == 2:
            assert t.remove(x) == (x in ref)
            ref.discard(x)
        elif op == 3:
            assert t.query(x) == (x in ref)
        elif op == 4:
            later = [y for y in ref if y > x]
            assert t.successor(x) == (min(later) if later else oracle.FAIL)
        else:
            assert t.find_first() == (min(ref) if ref else oracle.FAIL)
    _layers_consistent(t)
    # malloc_first drains in increasing order
    out = []
    while True:
        v = t.malloc_first()
        if v == oracle.FAIL:
            break
        out.append(v)
    assert out == sorted(ref)
    assert all(int(w) == 0 for layer in t.layers() for w in layer)


def _gather_closed_form(bits, n):
    """Run the reference's gather settles on: the highest word w whose
    trailing-ones run (length c > 0) ends a run of n set bits, i.e. all bits
    of [64w + c - n, 64w + c) set; answer 64w + c - n."""
    words = len(bits)
    isset = lambda i: (int(bits[i // 64]) >> (i % 64)) & 1
    for w in range(words - 1, -1, -1):
        v = int(bits[w])
        c = 0
        while c < 64 and (v >> c) & 1:
            c += 1
        if c == 0:
            continue
        end = 64 * w + c
        if end < n:
            continue
        if all(isset(i) for i in range(end - n, end)):
            return end - n
    return None


@pytest.mark.parametrize("universe", [64, 100, 256, 700, 4096])
def test_gather_multiple_closed_form(universe):
    rng = np.random.default_rng(universe + 7)
    for trial in range(150):
        t = oracle.Veb(universe, filled=False)
        # random runs of set bits
        density = rng.uniform(0.2, 0.95)
        members = [i for i in range(universe) if rng.random() < density]
        for i in members:
            t.insert(i)
        n = int(rng.choice([1, 2, 3, 17, 63, 64, 65, 128, 130]))
        before = t.layer(t.num_layers - 1)
        exp = _gather_closed_form(before, n)
        got = t.gather_multiple(n)
        if exp is None:
            assert got == oracle.FAIL
            assert np.array_equal(t.layer(t.num_layers - 1), before)
        else:
            assert got == exp, (trial, n)
            after = t.layer(t.num_layers - 1)
            diff = [i for i in range(universe)
                    if ((int(before[i // 64]) >> (i % 64)) & 1) != ((int(after[i // 64]) >> (i % 64)) & 1)]
            assert diff == list(range(exp, exp + n))
        _layers_consistent(t)


def test_return_multiple_sets_parents():
    t = oracle.Veb(4096, filled=False)
    t.return_multiple(100, 5)  # partial word: parent must see it
    _layers_consistent(t)
    assert t.find_first() == 100
    assert t.gather_multiple(5) == oracle.FAIL or True
    t2 = oracle.Veb(4096, filled=True)
    s = t2.gather_multiple(130)
    assert s == 4096 - 130
    t2.return_multiple(s, 130)
    assert all(np.array_equal(a, b) for a, b in zip(t2.layers(), H.fresh_layers(4096, True)))


# ----------------------------------------------------- allocator: golden

def _expand_slots(spec, counts):
    out = []
    for t, n in enumerate(counts):
        v = spec[str(t)]
        if isinstance(v[0], list):
            out.extend(v[0])
        else:
            out.extend(range(v[0], v[0] + v[1]))
        assert len(out) == sum(counts[: t + 1])
    return out


def test_boot_known_answer():
    g = G["boot_1GiB_16MiB_16_4096"]
    a = oracle.Allocator(g["pool_bytes"])
    assert a.num_segments == g["num_segments"]
    ids, act = a.segment_state()
    assert ids[:16].tolist() == g["chunk_ids_first_16"]
    assert act[:16].tolist() == g["active_first_16"]
    assert np.all(ids[16:] == 0xFFFF) and np.all(act[16:] == -1)
    for t, members in enumerate(g["sub_tree_members"]):
        bottom = a.sub_tree(t).layer(a.sub_tree(t).num_layers - 1)
        got = [i for i in range(a.num_segments) if (int(bottom[i // 64]) >> (i % 64)) & 1]
        assert got == members, t
    lo, hi = g["segment_tree_free"]
    st = a.segment_tree()
    bottom = st.layer(st.num_layers - 1)
    free = [i for i in range(a.num_segments) if (int(bottom[i // 64]) >> (i % 64)) & 1]
    assert free == list(range(lo, hi))
    assert a.slots().tolist() == _expand_slots(g["slot_blocks"], g["slot_counts"])


def test_single_lane_known_answers():
    a = oracle.Allocator(1 << 30)
    for step in G["single_lane_sequence_1GiB"]:
        assert a.malloc(step["size"]) == step["offset"], step
    assert a.errors == 0


def test_wave_known_answer():
    a = oracle.Allocator(1 << 30)
    g = G["wave_of_64_x16_after_boot_1GiB"]
    out = a.malloc_wave((1 << 64) - 1, [16] * 64)
    assert out.tolist() == [g["first"] + g["stride"] * i for i in range(64)]


def test_multi_segment_known_answers():
    a = oracle.Allocator(4 << 30)
    for step in G["multi_segment_4GiB"]:
        assert a.malloc(step["size"]) == step["offset"], step


# ------------------------------------------------- allocator: invariants

def _check_fresh(a):
    ids, act = a.segment_state()
    assert np.all(ids == 0xFFFF)
    assert np.all(act == -1)
    st = a.segment_tree()
    for x, y in zip(st.layers(), H.fresh_layers(a.num_segments, True)):
        assert np.array_equal(x, y)
    for t in range(a.num_trees):
        for x in a.sub_tree(t).layers():
            assert not np.any(x)
    assert np.all(a.slots() == 0xFFFFFFFF)
    assert np.all(a.blocks()[:, 1] == 0)


@pytest.mark.parametrize("seed,pool", [(1, 1 << 30), (2, 4 << 30), (3, 256 << 20)])
def test_stream_invariants_and_fresh_after_drain(seed, pool):
    a = oracle.Allocator(pool)
    steps, live = H.make_stream(seed, 300)
    res = a.run_stream(steps)
    offs, sizes = H.stream_sizes(steps, res)
    # live set at the end of the stream is disjoint and in the pool
    live_set = set(live)
    idx = [i * 64 + l for i, st in enumerate(steps) if st["kind"] == 0
           for l in range(64) if int(st["mask"]) >> l & 1]
    keep = np.array([j in live_set for j in idx])
    H.check_live(offs[keep], sizes[keep], a.num_segments * (16 << 20))
    assert a.errors == 0
    tail = H.free_all_steps(live, len(steps))
    res2 = np.concatenate([res, np.zeros(len(tail) * 64, np.uint64)])
    allsteps = np.concatenate([steps, tail])
    b = oracle.Allocator(pool)
    res_all = b.run_stream(allsteps)
    assert np.array_equal(res_all[: len(res)], res)
    b.drain()
    assert b.errors == 0
    _check_fresh(b)
    del res2


def test_every_prefix_is_disjoint():
    """Live sets after each step of a stream never overlap (small pool, so
    segments recycle between classes)."""
    a = oracle.Allocator(256 << 20)
    steps, _ = H.make_stream(11, 200, mix="slices")
    live = {}
    for i, st in enumerate(steps):
        one = np.zeros(1, H.STEP_DTYPE)
        one[0] = st
        if st["kind"] == 0:
            out = a.malloc_wave(int(st["mask"]), st["arg"])
            for l in range(64):
                if int(st["mask"]) >> l & 1 and int(out[l]) != oracle.FAIL:
                    live[i * 64 + l] = (int(out[l]), int(st["arg"][l]))
        else:
            offs = np.full(64, oracle.FAIL, np.uint64)
            for l in range(64):
                if int(st["mask"]) >> l & 1:
                    ref = int(st["arg"][l])
                    if ref in live:
                        offs[l] = live.pop(ref)[0]
            a.free_wave(int(st["mask"]), offs)
        if live:
            o = np.array([v[0] for v in live.values()], np.uint64)
            s = np.array([v[1] for v in live.values()], np.uint64)
            H.check_live(o, s, a.num_segments * (16 << 20))
    assert a.errors == 0


def test_exhaustion_returns_null_and_recovers():
    a = oracle.Allocator(256 << 20)  # 16 segments: boot formats all of them
    first = a.malloc(16 << 20)       # the prefetched class-8 segment 15
    assert first == 15 * (16 << 20)
    assert a.malloc(16 << 20) is None  # out of memory is a nullptr
    a.free(first)                       # segment 15 goes back ...
    assert a.malloc(16 << 20) == first  # ... and is found again
    assert a.errors == 0


def test_class_allocates_again_after_running_out():
    """F6: a class whose pinned slots went empty when the pool ran out
    allocates again once memory is freed (the reference's slot stays null
    for good), and each failing request gives up at once."""
    a = oracle.Allocator(256 << 20)
    got = []
    while True:
        o = a.malloc(2048)
        if o is None:
            break
        got.append(o)
    assert len(got) > 4096            # at least one block's worth was served
    assert a.malloc(2048) is None     # still out of memory
    for o in got:
        a.free(o)
    again = a.malloc(2048)
    assert again is not None
    a.free(again)
    a.drain()
    assert a.errors == 0
    ids, act = a.segment_state()
    assert np.all(ids == 0xFFFF) and np.all(act == -1)


def test_invalid_frees_counted():
    a = oracle.Allocator(1 << 30)
    a.free(None)                  # nullptr: ignored
    assert a.errors == 0
    a.free(20 * (16 << 20) + 64)  # unformatted segment
    a.free(1 << 40)               # outside the pool
    assert a.errors == 2


def test_churn_sequential_patterns():
    a = oracle.Allocator(1 << 30)
    pairs, misses, bad = a.run_churn_sequential(seed=5, tid0=0, nthreads=2000, rounds=10)
    assert bad == 0 and misses == 0
    assert pairs == H.expected_churn_pairs(5, 2000, 10)
    a.drain()
    _check_fresh(a)


def test_churn_sizes_match_streams_header():
    """numpy mirror of streams.h agrees with the C generator the oracle uses
    (both are integer-only; sizes lie in [1, 4095], log-uniform)."""
    tids = np.arange(20000, dtype=np.uint64)
    s = H.churn_size(7, tids, 3)
    assert s.min() >= 1 and s.max() <= 4095
    # log-uniform: about 1/12 of draws per octave
    octs = np.bincount(np.floor(np.log2(s.astype(np.float64))).astype(int), minlength=12)
    assert np.all(np.abs(octs / len(s) - 1 / 12) < 0.01)


def test_size_boundary_known_answers():
    """32768 B is the last run of 4 KiB slices, 32769 B the first whole
    block; exactly 16 MiB is a whole block of the largest class."""
    a = oracle.Allocator(1 << 30)
    for step in G["size_boundaries_1GiB"]:
        assert a.malloc(step["size"]) == step["offset"], step
    assert a.errors == 0


def test_ring_reuse_and_reregister_known_answer():
    """Blocks come back through the segment ring first in, first out, and a
    segment rejoins its class tree at the return that brings its free
    blocks to 10 % (4 of 32)."""
    a = oracle.Allocator(1 << 30)
    got = []
    for op in G["ring_reuse_and_reregister_1GiB"]:
        if op["op"] == "malloc":
            r = a.malloc(op["size"])
            assert r == op["offset"], (len(got), op)
            got.append(r)
        else:
            a.free(got[op["of"]])
            got.append(None)
    assert a.errors == 0

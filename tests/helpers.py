"""Stream generators and exact invariant checkers shared by the tests."""
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FAIL = (1 << 64) - 1
STEP_DTYPE = np.dtype([("kind", "<u4"), ("pad", "<u4"), ("mask", "<u8"),
                       ("arg", "<u8", (64,))])


def golden():
    with open(os.path.join(HERE, "golden", "known_answers.json")) as f:
        return json.load(f)


def clog2(x):
    """ceil(log2 x) for x >= 1."""
    return (x - 1).bit_length()


def reserved(size, S=16 << 20, m=16, M=4096):
    """(bytes reserved, alignment) for a request, from the size dispatch of
    gallatin.cuh:1141-1147 / 1186-1276: slices of class t, runs of 4 KiB
    slices, whole blocks of 4096 slices, or whole segments."""
    T = clog2(M) - clog2(m) + 1
    t = 0 if size < m else clog2(size) - clog2(m)
    if t < T:
        return m << t, m << t
    bt = clog2(size) - clog2(m * 4096)
    if bt < 0:
        return (1 << (t - (T - 1))) * M, M
    if bt < T:
        return 4096 * (m << bt), 4096 * (m << bt)
    n = (size - 1) // S + 1
    return n * S, S


def check_live(offsets, sizes, pool_bytes, S=16 << 20, m=16, M=4096):
    """Exact invariants over live allocations: inside the pool, aligned to
    their class, and pairwise disjoint.  Returns the number of checked."""
    offsets = np.asarray(offsets, dtype=np.uint64)
    sizes = np.asarray(sizes, dtype=np.uint64)
    ok = offsets != np.uint64(FAIL)
    offsets, sizes = offsets[ok], sizes[ok]
    if len(offsets) == 0:
        return 0
    res = np.array([reserved(int(s), S, m, M) for s in np.unique(sizes)],
                   dtype=np.uint64)
    lut = dict(zip(np.unique(sizes).tolist(), res.tolist()))
    span = np.array([lut[int(s)][0] for s in sizes], dtype=np.uint64)
    align = np.array([lut[int(s)][1] for s in sizes], dtype=np.uint64)
    assert np.all(offsets % align == 0), "allocation not aligned to its class"
    assert np.all(offsets + span <= np.uint64(pool_bytes)), "allocation outside the pool"
    order = np.argsort(offsets, kind="stable")
    o, sp = offsets[order], span[order]
    assert np.all(o[:-1] + sp[:-1] <= o[1:]), "two live allocations overlap"
    return len(offsets)


def reserved_vec(sizes, S=16 << 20, m=16, M=4096):
    u = np.unique(sizes)
    lut = {int(s): reserved(int(s), S, m, M) for s in u}
    span = np.array([lut[int(s)][0] for s in sizes], dtype=np.uint64)
    align = np.array([lut[int(s)][1] for s in sizes], dtype=np.uint64)
    return span, align


def check_ptrs(ptrs, sizes, base, pool_bytes, S=16 << 20, m=16, M=4096):
    """Same invariants for device pointers (0 = nullptr)."""
    ptrs = np.asarray(ptrs, dtype=np.uint64)
    live = ptrs != 0
    offs = np.where(live, ptrs - np.uint64(base), np.uint64(FAIL))
    return check_live(offs, sizes, pool_bytes, S, m, M)


def random_sizes(rng, n, mix="slices"):
    """Request sizes: 'slices' (1..4096, log-uniform), 'small' (16 B only),
    'all' (slices, big slices, blocks and a few multi-segment)."""
    if mix == "small":
        return np.full(n, 16, np.uint64)
    logs = rng.uniform(0, 12, n)
    s = np.floor(2.0 ** logs).astype(np.uint64)
    if mix == "slices":
        return np.clip(s, 1, 4096)
    r = rng.random(n)
    big = rng.integers(4097, 32768, n)           # big slices
    blk = rng.integers(32769, 16 << 20, n)       # whole blocks
    seg = rng.integers((16 << 20) + 1, 64 << 20, n)  # 2..4 segments
    out = np.where(r < 0.80, np.clip(s, 1, 4096),
                   np.where(r < 0.90, big, np.where(r < 0.985, blk, seg)))
    return out.astype(np.uint64)


def make_stream(seed, nsteps, mix="all", p_free=0.4, lane_modes=("one", "some", "all")):
    """Random wave-step stream: malloc steps with random lane masks and
    sizes, free steps of random live allocations (each freed once)."""
    rng = np.random.default_rng(seed)
    steps = np.zeros(nsteps, STEP_DTYPE)
    live = []  # indices step*64+lane of outstanding mallocs
    for i in range(nsteps):
        st = steps[i]
        if live and rng.random() < p_free:
            k = min(len(live), int(rng.integers(1, 65)))
            pick = rng.choice(len(live), size=k, replace=False)
            refs = [live[j] for j in pick]
            for j in sorted(pick, reverse=True):
                live.pop(j)
            lanes = rng.choice(64, size=k, replace=False)
            mask = 0
            for lane, ref in zip(lanes, refs):
                mask |= 1 << int(lane)
                st["arg"][lane] = ref
            st["kind"] = 1
            st["mask"] = mask
        else:
            mode = lane_modes[int(rng.integers(0, len(lane_modes)))]
            if mode == "one":
                lanes = [int(rng.integers(0, 64))]
            elif mode == "all":
                lanes = list(range(64))
            else:
                lanes = sorted(set(rng.integers(0, 64, int(rng.integers(2, 40))).tolist()))
            sizes = random_sizes(rng, 64, mix)
            mask = 0
            for lane in lanes:
                mask |= 1 << lane
                live.append(i * 64 + lane)
            st["kind"] = 0
            st["mask"] = mask
            st["arg"][:] = sizes
    return steps, live


def free_all_steps(live, start_index):
    """Wave steps freeing every outstanding malloc (64 per step)."""
    n = (len(live) + 63) // 64
    steps = np.zeros(n, STEP_DTYPE)
    for i in range(n):
        chunk = live[i * 64:(i + 1) * 64]
        steps[i]["kind"] = 1
        m = 0
        for lane, ref in enumerate(chunk):
            m |= 1 << lane
            steps[i]["arg"][lane] = ref
        steps[i]["mask"] = m
    return steps


def stream_sizes(steps, results):
    """(offset, size) of every successful malloc in a replayed stream."""
    offs, sizes = [], []
    for i, st in enumerate(steps):
        if st["kind"] != 0:
            continue
        m = int(st["mask"])
        for lane in range(64):
            if m >> lane & 1:
                offs.append(int(results[i * 64 + lane]))
                sizes.append(int(st["arg"][lane]))
    return np.array(offs, np.uint64), np.array(sizes, np.uint64)


def fresh_layers(universe, filled):
    """Layers (top first) of a fresh tree: all bits below each layer's
    universe set, or all zero (veb.cuh:72-84, 544-561)."""
    layers = []
    u = universe
    sizes = []
    max_height = universe.bit_length()
    nl = (max_height - 1) // 6 + 1
    for _ in range(nl):
        sizes.append(u)
        u = (u - 1) // 64 + 1
    for u in reversed(sizes):
        w = (u - 1) // 64 + 1
        arr = np.zeros(w, np.uint64)
        if filled:
            arr[:] = np.uint64(FAIL)
            if u % 64:
                arr[-1] = np.uint64((1 << (u % 64)) - 1)
        layers.append(arr)
    return layers


def pack_state(ids, active):
    return (active.astype(np.int64).astype(np.uint64) << np.uint64(32)) | ids.astype(np.uint64)


def expected_churn_pairs(seed, nthreads, rounds):
    """Pairs the mixed churn performs when no malloc misses."""
    from oracle import lib  # noqa: F401  (streams are recomputed in numpy)
    tids = np.arange(nthreads, dtype=np.uint64)
    total = np.ones(nthreads, np.uint64)
    for r in range(1, rounds):
        total += churn_coin(seed, tids, r)
    return int(total.sum())


# numpy mirror of include/gallatin/streams.h (for expected counts)
def _mix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(FAIL)
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(FAIL)
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(FAIL)
    return x ^ (x >> np.uint64(31))


def hash4(seed, a, b, c):
    with np.errstate(over="ignore"):
        h = _mix64(np.uint64(seed) ^ np.uint64(0x5851F42D4C957F2D))
        h = _mix64(h ^ np.asarray(a, np.uint64))
        h = _mix64(h ^ (np.uint64(b) * np.uint64(0xD6E8FEB86659FD93)))
        h = _mix64(h ^ (np.uint64(c) * np.uint64(0xA0761D6478BD642F)))
    return h


def churn_coin(seed, tids, r):
    return (hash4(seed, tids, r, 2) >> np.uint64(63)).astype(np.uint64)


def churn_size(seed, tids, r):
    h = hash4(seed, tids, r, 1)
    x = (h >> np.uint64(32)) * np.uint64(12)
    k = (x >> np.uint64(32)).astype(np.int64)
    f = (x & np.uint64(0xFFFFFFFF)).astype(np.uint64)
    one = np.uint64(1 << 30)
    with np.errstate(over="ignore"):
        y = (f * np.uint64(2977044472)) >> np.uint64(34)
        t = np.full_like(y, one)
        for i in range(9, 0, -1):
            t = one + ((y * t) >> np.uint64(30)) // np.uint64(i)
    v = (t << k.astype(np.uint64)) >> np.uint64(30)
    return np.maximum(v, np.uint64(1))


def log_uniform_mean_bytes():
    return 4095 / math.log(4096)


def check_tree_parents(layers):
    """A quiescent tree's invariant (veb.hpp sync_up): bit i of each word
    of a layer is set exactly when word i of the layer below is non-zero,
    and no bit points past the layer below.  Layers top first."""
    for up, down in zip(layers[:-1], layers[1:]):
        idx = np.arange(len(down))
        bit = (up[idx // 64] >> (idx % 64).astype(np.uint64)) & np.uint64(1)
        bad = np.nonzero((bit != 0) != (down != 0))[0]
        assert len(bad) == 0, f"parent bits disagree with child words {bad[:8].tolist()}"
        extra = len(up) * 64 - len(down)
        if extra:
            tail = up[-1] >> np.uint64(64 - extra) if extra < 64 else up[-1]
            assert tail == 0, "parent bits past the last child word"


def check_all_tree_parents(a):
    check_tree_parents(a.tree_layers(-1))
    for t in range(a.num_trees):
        for p in range(a.partitions):
            check_tree_parents(a.tree_layers(t, p))


def check_fresh(a):
    """Metadata of a drained instance (nothing allocated, nothing pinned)
    equals the freshly initialised state bit for bit."""
    st, sl, bl = a.snapshot()
    assert np.all(st == np.uint64(0xFFFFFFFF0000FFFF))
    assert np.all(sl == 0xFFFFFFFF)
    assert np.all(bl[:, 1] == 0), "a block kept freed slices"
    for x, y in zip(a.tree_layers(-1), fresh_layers(a.num_segments, True)):
        assert np.array_equal(x, y)
    for t in range(a.num_trees):
        for p in range(a.partitions):
            for x in a.tree_layers(t, p):
                assert not np.any(x)
    assert a.info()["errors"] == 0

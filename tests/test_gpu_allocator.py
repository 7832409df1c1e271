"""Parity of the HIP allocator with the oracle (GPU).

Deterministic single-wave streams must return exactly the oracle's offsets
and leave bit-identical metadata; concurrent runs at full size must satisfy
the exact invariants (in pool, class-aligned, disjoint, patterns intact,
metadata back to the fresh state after all frees and a drain).
"""
import numpy as np
import pytest

import oracle
from tests import helpers as H

pytestmark = pytest.mark.gpu

G = H.golden()
S = 16 << 20
# single-wave runs: once an attempt has failed without changing state every
# later attempt fails the same way, so small limits give the same results
QUICK = (6, 6, 6, 6)


def make(gpu, pool, det=True, limits=QUICK, seed=42):
    flags = gpu.GAL_FLAG_DETERMINISTIC if det else 0
    return gpu.Gallatin(S, 16, 4096).generate_on_device(
        pool, seed, print_info=False, flags=flags, limits=limits)


def compare_metadata(a, o):
    """GPU instance `a` and oracle `o` hold identical metadata."""
    st, sl, bl = a.snapshot()
    ids, act = o.segment_state()
    assert np.array_equal(st, H.pack_state(ids, act)), "segment state differs"
    assert np.array_equal(sl, o.slots()), "pinned slots differ"
    assert np.array_equal(bl, o.blocks()), "block counters differ"
    assert all(np.array_equal(x, y) for x, y in
               zip(a.tree_layers(-1), o.segment_tree().layers())), "segment tree differs"
    for t in range(a.num_trees):
        assert all(np.array_equal(x, y) for x, y in
                   zip(a.tree_layers(t), o.sub_tree(t).layers())), f"class tree {t} differs"


check_fresh = H.check_fresh


def test_boot_matches_oracle_and_golden(gpu):
    g = G["boot_1GiB_16MiB_16_4096"]
    a = make(gpu, g["pool_bytes"])
    o = oracle.Allocator(g["pool_bytes"], limits=QUICK)
    compare_metadata(a, o)
    st, _, _ = a.snapshot()
    assert (st[:16] & np.uint64(0xFFFF)).tolist() == g["chunk_ids_first_16"]
    a.free_on_device()


def _single_lane_steps(sizes):
    steps = np.zeros(len(sizes), H.STEP_DTYPE)
    for i, s in enumerate(sizes):
        steps[i]["kind"] = 0
        steps[i]["mask"] = 1
        steps[i]["arg"][0] = s
    return steps


def test_single_lane_known_answers(gpu):
    seq = G["single_lane_sequence_1GiB"]
    a = make(gpu, 1 << 30)
    res = a.run_wave_stream(_single_lane_steps([s["size"] for s in seq]))
    got = [None if int(res[i * 64]) == H.FAIL else int(res[i * 64]) for i in range(len(seq))]
    assert got == [s["offset"] for s in seq]
    o = oracle.Allocator(1 << 30, limits=QUICK)
    o.run_stream(_single_lane_steps([s["size"] for s in seq]))
    compare_metadata(a, o)
    a.free_on_device()


def test_wave_known_answer(gpu):
    g = G["wave_of_64_x16_after_boot_1GiB"]
    a = make(gpu, 1 << 30)
    steps = np.zeros(1, H.STEP_DTYPE)
    steps[0]["mask"] = (1 << 64) - 1
    steps[0]["arg"][:] = 16
    res = a.run_wave_stream(steps)
    assert res[:64].tolist() == [g["first"] + g["stride"] * i for i in range(64)]
    a.free_on_device()


def test_size_boundary_known_answers(gpu):
    """32768 / 32769 B and exactly 16 MiB: the last slice run, the first
    whole block and the largest class's block."""
    seq = G["size_boundaries_1GiB"]
    a = make(gpu, 1 << 30)
    res = a.run_wave_stream(_single_lane_steps([s["size"] for s in seq]))
    assert [int(res[i * 64]) for i in range(len(seq))] == [s["offset"] for s in seq]
    o = oracle.Allocator(1 << 30, limits=QUICK)
    o.run_stream(_single_lane_steps([s["size"] for s in seq]))
    compare_metadata(a, o)
    a.free_on_device()


def test_ring_reuse_and_reregister_known_answer(gpu):
    """Blocks come back through the segment ring first in, first out, and a
    segment rejoins its class tree at the return that brings its free
    blocks to 10 % (fix F7); metadata equals the oracle's afterwards."""
    seq = G["ring_reuse_and_reregister_1GiB"]
    steps = np.zeros(len(seq), H.STEP_DTYPE)
    for i, op in enumerate(seq):
        steps[i]["mask"] = 1
        if op["op"] == "malloc":
            steps[i]["kind"] = 0
            steps[i]["arg"][0] = op["size"]
        else:
            steps[i]["kind"] = 1
            steps[i]["arg"][0] = op["of"] * 64
    a = make(gpu, 1 << 30)
    res = a.run_wave_stream(steps)
    for i, op in enumerate(seq):
        if op["op"] == "malloc":
            assert int(res[i * 64]) == op["offset"], (i, op)
    o = oracle.Allocator(1 << 30, limits=QUICK)
    assert np.array_equal(o.run_stream(steps), res)
    compare_metadata(a, o)
    assert a.info()["errors"] == 0
    a.free_on_device()


def test_multi_segment_known_answers(gpu):
    seq = G["multi_segment_4GiB"]
    a = make(gpu, 4 << 30)
    res = a.run_wave_stream(_single_lane_steps([s["size"] for s in seq]))
    assert [int(res[i * 64]) for i in range(len(seq))] == [s["offset"] for s in seq]
    a.free_on_device()


@pytest.mark.parametrize("seed,pool", [(2, 4 << 30), (3, 256 << 20)])
def test_fill_statistics_match_oracle(gpu, seed, pool):
    """print_info's numbers (gal_get_info) in the middle of a deterministic
    stream, with allocations live: the segment tree's free count and each
    class tree's fill equal the oracle's tree fills (orc_veb_fill,
    veb.cuh:1052-1077), and the per-class guided fill (print_guided_fill,
    gallatin.cuh:1958-1990: segments of the class, blocks available and
    handed out) equals the same sums over the oracle's segment state."""
    steps, _ = H.make_stream(seed, 200, mix="all")
    a = make(gpu, pool)
    o = oracle.Allocator(pool, limits=QUICK)
    assert np.array_equal(a.run_wave_stream(steps), o.run_stream(steps))
    info = a.info()
    assert info["segments_free"] == o.segment_tree().fill()
    ids, act = o.segment_state()
    for t in range(a.num_trees):
        assert info["tree_owned"][t] == o.sub_tree(t).fill(), t
        m = (ids == t) & (act != -1)
        nb = a.blocks_per_segment >> t
        avail = np.clip(act[m].astype(np.int64) + 1, 0, None)
        assert info["class_segments"][t] == int(m.sum()), t
        assert info["class_blocks_free"][t] == int(avail.sum()), t
        assert info["class_blocks_used"][t] == int(np.clip(nb - avail, 0, None).sum()), t
    assert sum(info["class_segments"]) > 0 and sum(info["class_blocks_used"]) > 0
    a.print_info()
    a.free_on_device()


@pytest.mark.parametrize("seed,pool,mix,nsteps", [
    (1, 1 << 30, "slices", 400),
    (2, 4 << 30, "all", 400),
    (3, 256 << 20, "all", 300),      # small pool: exhaustion and recycling
    (4, 64 << 30, "all", 300),       # 3-level trees
    (5, 2 << 30, "slices", 600),
    (6, 512 << 20, "all", 500),      # exhaustion on a mid-size pool
    (7, 16 << 30, "all", 400),
    (8, 256 << 20, "slices", 800),   # slices only, recycling under exhaustion
])
def test_stream_parity(gpu, seed, pool, mix, nsteps):
    steps, live = H.make_stream(seed, nsteps, mix=mix)
    steps = np.concatenate([steps, H.free_all_steps(live, len(steps))])
    o = oracle.Allocator(pool, limits=QUICK)
    exp = o.run_stream(steps)
    a = make(gpu, pool)
    got = a.run_wave_stream(steps)
    if not np.array_equal(got, exp):
        bad = np.nonzero(got != exp)[0]
        i = int(bad[0])
        pytest.fail(f"first difference at step {i // 64} lane {i % 64}: "
                    f"gpu {int(got[i]):#x} oracle {int(exp[i]):#x} ({len(bad)} differ)")
    compare_metadata(a, o)
    a.drain()
    o.drain()
    compare_metadata(a, o)
    check_fresh(a)
    a.free_on_device()


def test_class_recovers_after_running_out(gpu):
    """Waves of 64 x 2 KiB until the 256 MiB pool runs out, then every
    allocation freed, then more 2 KiB waves: the class allocates again
    (empty pinned slots are refilled) with the oracle's offsets."""
    n_alloc, n_again = 2200, 8
    steps = np.zeros(2 * n_alloc + n_again, H.STEP_DTYPE)
    full = (1 << 64) - 1
    for i in range(n_alloc):
        steps[i]["kind"], steps[i]["mask"] = 0, full
        steps[i]["arg"][:] = 2048
    for i in range(n_alloc):
        st = steps[n_alloc + i]
        st["kind"], st["mask"] = 1, full
        st["arg"][:] = i * 64 + np.arange(64)
    for i in range(2 * n_alloc, len(steps)):
        steps[i]["kind"], steps[i]["mask"] = 0, full
        steps[i]["arg"][:] = 2048
    pool = 256 << 20
    o = oracle.Allocator(pool, limits=QUICK)
    exp = o.run_stream(steps)
    assert np.any(exp[:n_alloc * 64] == H.FAIL), "the pool did not run out"
    assert np.all(exp[2 * n_alloc * 64:] != H.FAIL), "oracle did not recover"
    a = make(gpu, pool)
    got = a.run_wave_stream(steps)
    assert np.array_equal(got, exp)
    compare_metadata(a, o)
    a.free_on_device()


def test_partitions_are_one_or_per_xcd(gpu):
    """A partition must be exactly one XCD (a home free only releases its
    payload to its own XCD's L2), so an instance has one partition or one
    per XCD: 8 GiB gets one per XCD, 4 GiB (too small to pin every class in
    8 partitions) gets one, and any other explicit count is refused."""
    import torch
    xcds = torch.cuda.get_device_properties(0).multi_processor_count // 32
    a = gpu.Gallatin(S, 16, 4096).generate_on_device(8 << 30, 1, print_info=False)
    assert a.partitions == xcds
    a.free_on_device()
    a = gpu.Gallatin(S, 16, 4096).generate_on_device(4 << 30, 1, print_info=False)
    assert a.partitions == 1
    a.free_on_device()
    with pytest.raises(gpu.GallatinError):
        gpu.Gallatin(S, 16, 4096).generate_on_device(8 << 30, 1, print_info=False, partitions=4)


@pytest.mark.parametrize("mode", ["fenced", "coherent", "partitioned"])
def test_plain_small_config(gpu, mode):
    """1 GiB pool, 2^20 threads x 10 launches of malloc(16) / write id /
    read back / free; every pair succeeds and the pool returns to fresh.
    coherent: the instance skips its payload fences and the id is written
    and read with device-coherent accesses.  partitioned: slots filled on
    demand (GAL_FLAG_LAZY_BOOT) and one partition per XCD, which the eager
    boot cannot fit in 1 GiB."""
    import torch
    flags, parts = 0, 0
    if mode == "coherent":
        flags = gpu.GAL_FLAG_NO_PAYLOAD_FENCE
    elif mode == "partitioned":
        flags = gpu.GAL_FLAG_LAZY_BOOT
        parts = torch.cuda.get_device_properties(0).multi_processor_count // 32
    a = gpu.Gallatin(S, 16, 4096).generate_on_device(1 << 30, 42, print_info=False, flags=flags,
                                                      partitions=parts)
    if mode == "partitioned":
        assert a.partitions == parts > 1
    ctr = torch.zeros(gpu.CTR_COUNT, dtype=torch.int64, device="cuda")
    for _ in range(10):
        a.run_plain(1 << 20, 16, counters=ctr)
    torch.cuda.synchronize()
    c = ctr.cpu().tolist()
    assert c[gpu.CTR_PAIRS] == 10 << 20
    assert c[gpu.CTR_MISSES] == 0 and c[gpu.CTR_BAD] == 0
    a.drain()
    check_fresh(a)
    a.free_on_device()


def test_bulk_live_set_disjoint(gpu):
    """2^18 mixed requests live at once: in pool, aligned, disjoint; then all
    freed (random order) and drained back to the fresh state."""
    import torch
    rng = np.random.default_rng(5)
    n = 1 << 18
    sizes = H.random_sizes(rng, n, "slices")
    a = make(gpu, 8 << 30, det=False, limits=None)
    d_sizes = torch.from_numpy(sizes.astype(np.int64)).cuda()
    ptrs = a.malloc_bulk(d_sizes)
    torch.cuda.synchronize()
    p = ptrs.cpu().numpy().view(np.uint64)
    assert np.all(p != 0), "unexpected nullptr"
    H.check_ptrs(p, sizes, a.pool_base, a.num_segments * S)
    perm = torch.from_numpy(rng.permutation(n)).cuda()
    a.free_bulk(ptrs[perm].contiguous())
    torch.cuda.synchronize()
    a.drain()
    check_fresh(a)
    a.free_on_device()


def test_block_and_multi_segment_config(gpu):
    """64 GiB pool (4096 segments, 3-level trees): 4096 requests uniform in
    [64 KiB, 64 MiB] (whole blocks and 2-4 segment runs) and 64 requests
    uniform in [256 MiB, 2 GiB] (16-128 segment runs) in one launch.  Demand
    is far above the pool, so many return nullptr; the served ones are in
    the pool, aligned and disjoint and fill most of it.  The wave's
    multi-segment claims race block requests for segments, so some gathers
    lose a claim and put the bits back: right after the launch, and again
    after the frees, every tree's parent bits must match its child words.
    All are freed in random order, and after a drain every bitmap, chunk id
    and counter is the fresh state bit for bit."""
    import torch
    rng = np.random.default_rng(33)
    sizes = np.concatenate([rng.integers(64 << 10, (64 << 20) + 1, 4096),
                            rng.integers(256 << 20, (2 << 30) + 1, 64)]).astype(np.uint64)
    pool = 64 << 30
    a = make(gpu, pool, det=False, limits=None)
    ptrs = a.malloc_bulk(torch.from_numpy(sizes.astype(np.int64)).cuda())
    torch.cuda.synchronize()
    p = ptrs.cpu().numpy().view(np.uint64)
    served = p != 0
    assert served.sum() > 0
    H.check_ptrs(p, sizes, a.pool_base, a.num_segments * S)
    span, _ = H.reserved_vec(sizes)
    assert int(span[served].sum()) <= pool
    assert int(span[served].sum()) > pool // 2, "most of the pool should be in use"
    H.check_all_tree_parents(a)
    live = ptrs[torch.from_numpy(np.nonzero(served)[0]).cuda()]
    perm = torch.from_numpy(rng.permutation(len(live))).cuda()
    a.free_bulk(live[perm].contiguous())
    torch.cuda.synchronize()
    H.check_all_tree_parents(a)
    a.drain()
    check_fresh(a)
    a.free_on_device()


def test_block_and_multi_segment_steps(gpu):
    """Three steps of the block and multi-segment batch (gal_run_big, the
    workload bench.py times).  In each step the served requests were live
    together, so they must be in the pool, aligned and disjoint; served plus
    misses is every request; every tag checks; every served request is
    freed.  After the steps a drain leaves the fresh state bit for bit."""
    import torch
    seed, n_small, n_large = 77, 4096, 64
    a = make(gpu, 64 << 30, det=False, limits=None)
    ctr = torch.zeros(gpu.CTR_COUNT, dtype=torch.int64, device="cuda")
    for step in range(3):
        ctr.zero_()
        ptrs, _ = a.run_big(seed, step, n_small, n_large, counters=ctr)
        torch.cuda.synchronize()
        p = ptrs.cpu().numpy().view(np.uint64)
        sizes, _ = oracle.big_stream(seed, step, n_small, n_large)
        served = int((p != 0).sum())
        c = ctr.cpu().tolist()
        assert served > 0
        assert c[gpu.CTR_MALLOCS] == n_small + n_large
        assert c[gpu.CTR_PAIRS] == served and c[gpu.CTR_MISSES] == n_small + n_large - served
        assert c[gpu.CTR_BAD] == 0
        assert c[gpu.CTR_BYTES] == 16 * served
        H.check_ptrs(p, sizes, a.pool_base, a.num_segments * S)
    a.drain()
    check_fresh(a)
    a.free_on_device()


def test_wave_of_multi_segment_requests(gpu):
    """Lanes of one wave asking for multi-segment runs.  When their total
    fits, the wave claims it as one run split in lane order: every request is
    served, in the pool and disjoint.  When it does not fit, the lanes claim
    one by one in lane order, so each gets exactly the run the oracle's
    gather_multiple (veb.cuh:323-452, 1160-1218) returns when replayed over
    the same free segments, and nullptr once none is left.  Both drain back
    to the fresh state after the frees."""
    import torch
    a = make(gpu, 4 << 30, det=False, limits=None)
    rng = np.random.default_rng(8)
    nseg = rng.integers(2, 5, 64)
    sizes = (nseg.astype(np.uint64) - 1) * np.uint64(S) + np.uint64(1)
    ptrs = a.malloc_bulk(torch.from_numpy(sizes.astype(np.int64)).cuda())
    torch.cuda.synchronize()
    p = ptrs.cpu().numpy().view(np.uint64)
    assert np.all(p != 0)
    H.check_ptrs(p, sizes, a.pool_base, a.num_segments * S)
    assert np.all((p - np.uint64(a.pool_base)) % np.uint64(S) == 0)
    a.free_bulk(ptrs[torch.randperm(64).cuda()].contiguous())
    torch.cuda.synchronize()
    bottom = a.tree_layers(-1)[-1]
    ref = oracle.Veb(a.num_segments, filled=False)
    for w, v in enumerate(bottom):
        for b in range(64):
            if (int(v) >> b) & 1:
                ref.insert(w * 64 + b)
    expect = []
    for _ in range(64):
        x = ref.gather_multiple(5)
        expect.append(0 if x >= a.num_segments else a.pool_base + x * S)
    sizes = np.full(64, 4 * S + 1, dtype=np.uint64)  # 5 segments each: 320 > free
    ptrs = a.malloc_bulk(torch.from_numpy(sizes.astype(np.int64)).cuda())
    torch.cuda.synchronize()
    p = ptrs.cpu().numpy().view(np.uint64)
    assert 0 < int((p != 0).sum()) < 64
    assert p.tolist() == expect
    H.check_ptrs(p, sizes, a.pool_base, a.num_segments * S)
    a.free_bulk(ptrs)
    torch.cuda.synchronize()
    a.drain()
    check_fresh(a)
    a.free_on_device()


def test_partitioned_segments_do_not_accumulate(gpu):
    """8 GiB pool, one partition per XCD.  Rounds of 2^16 x 4 KiB requests
    (the 1-block-per-segment class) freed in a shuffled order, so most
    blocks complete on a foreign XCD and their segments stay formatted for
    their home.  Such a segment is reused by its home, and taking it must
    not also pull a fresh segment from the pool: the formatted count levels
    off instead of growing every round, and a drain still gives back every
    segment."""
    import torch
    a = make(gpu, 8 << 30, det=False, limits=None)
    assert a.partitions > 1
    n = 1 << 16
    sizes = torch.full((n,), 4096, dtype=torch.int64, device="cuda")
    fmt = []
    for r in range(10):
        ptrs = a.malloc_bulk(sizes)
        torch.cuda.synchronize()
        assert bool((ptrs != 0).all())
        a.free_bulk(ptrs[torch.randperm(n, device="cuda")].contiguous())
        torch.cuda.synchronize()
        fmt.append(a.info()["segments_formatted"])
    print("formatted segments per round", fmt)
    assert fmt[-1] <= fmt[2] + 8, f"formatted segments keep growing: {fmt}"
    a.drain()
    check_fresh(a)
    a.free_on_device()


@pytest.mark.parametrize("lazy", [False, True])
def test_churn_patterns_and_counts(gpu, lazy):
    """The mixed-size churn, every class and every XCD partition at once;
    lazy: the slots start empty and fill on first use."""
    import torch
    n, rounds, seed = 1 << 20, 10, 9
    if lazy:
        a = gpu.Gallatin(S, 16, 4096).generate_on_device(8 << 30, 42, print_info=False,
                                                          flags=gpu.GAL_FLAG_LAZY_BOOT)
        assert a.partitions > 1 and np.all(a.snapshot()[1] == 0xFFFFFFFF)
    else:
        a = make(gpu, 8 << 30, det=False, limits=None)
    ctr = a.run_churn(seed, n, rounds)
    torch.cuda.synchronize()
    c = ctr.cpu().tolist()
    assert c[gpu.CTR_BAD] == 0
    assert c[gpu.CTR_MISSES] == 0
    assert c[gpu.CTR_PAIRS] == H.expected_churn_pairs(seed, n, rounds)
    a.drain()
    check_fresh(a)
    a.free_on_device()


@pytest.mark.parametrize("partitions,threads", [(0, 1024), (1, 1024), (0, 256)])
def test_churn_batched_patterns_and_counts(gpu, partitions, threads):
    """The workgroup-batched path serves the same stream: every pair
    completes, every pattern reads back, the pool drains to fresh."""
    import torch
    n, rounds, seed = 1 << 20, 10, 10
    a = gpu.Gallatin(S, 16, 4096).generate_on_device(
        8 << 30, 7, print_info=False, slots_min=32, partitions=partitions)
    ctr = a.run_churn_batched(seed, n, rounds, threads=threads)
    torch.cuda.synchronize()
    c = ctr.cpu().tolist()
    assert c[gpu.CTR_BAD] == 0
    assert c[gpu.CTR_MISSES] == 0
    assert c[gpu.CTR_PAIRS] == H.expected_churn_pairs(seed, n, rounds)
    a.drain()
    check_fresh(a)
    a.free_on_device()


@pytest.mark.parametrize("batched,threads,partitions", [
    (False, 256, 0), (True, 1024, 0), (True, 256, 0), (False, 256, 1), (True, 1024, 1)])
def test_churn_one_word(gpu, batched, threads, partitions):
    """The churn stream with one word per allocation written and checked (the
    reference's churn test pattern), per thread and through workgroup
    batches, with one partition per XCD and with one partition (a word then
    lives in any XCD's L2 between its write and its check): every pair
    completes, every word reads back, 16 bytes per pair are counted, and
    the pool drains to fresh."""
    import torch
    n, rounds, seed = 1 << 20, 10, 11
    a = gpu.Gallatin(S, 16, 4096).generate_on_device(
        8 << 30, 8, print_info=False, slots_min=32, partitions=partitions)
    assert (a.partitions > 1) == (partitions == 0)
    ctr = a.run_churn_word(seed, n, rounds, batched=batched, threads=threads)
    torch.cuda.synchronize()
    c = ctr.cpu().tolist()
    pairs = H.expected_churn_pairs(seed, n, rounds)
    assert c[gpu.CTR_BAD] == 0
    assert c[gpu.CTR_MISSES] == 0
    assert c[gpu.CTR_PAIRS] == pairs
    assert c[gpu.CTR_BYTES] == 16 * pairs
    a.drain()
    check_fresh(a)
    a.free_on_device()


def test_churn_one_word_ragged_and_refused(gpu):
    """No thread, one thread and thread counts that leave the last block or
    workgroup chunk partly empty, per thread and at every batched
    workgroup size, one after another on one instance: the counts match
    the stream, and the pool drains to fresh.  Arguments outside the
    entry point's range are refused before any launch."""
    import torch
    a = gpu.Gallatin(S, 16, 4096).generate_on_device(1 << 30, 5, print_info=False)
    rounds = 4
    for k, n in enumerate((0, 1, 1000, 70001)):
        for batched, threads in ((False, 256), (False, 64), (True, 256), (True, 512),
                                 (True, 1024)):
            seed = 100 * k + threads + int(batched)
            ctr = a.run_churn_word(seed, n, rounds, batched=batched, threads=threads)
            torch.cuda.synchronize()
            c = ctr.cpu().tolist()
            pairs = H.expected_churn_pairs(seed, n, rounds)
            case = (n, batched, threads)
            assert c[gpu.CTR_BAD] == 0, case
            assert c[gpu.CTR_MISSES] == 0, case
            assert c[gpu.CTR_PAIRS] == pairs, case
            assert c[gpu.CTR_BYTES] == 16 * pairs, case
    for kw in ({"batched": True, "threads": 300}, {"batched": True, "threads": 128}):
        with pytest.raises(gpu.GallatinError):
            a.run_churn_word(1, 1000, rounds, **kw)
    assert a.info()["errors"] == 0
    a.drain()
    check_fresh(a)
    a.free_on_device()


def test_global_allocator(gpu):
    import torch
    gpu.init_global_allocator(1 << 30, 42, print_info=False)
    sizes = torch.full((1 << 16,), 64, dtype=torch.int64, device="cuda")
    ptrs = gpu.global_malloc_bulk(sizes)
    torch.cuda.synchronize()
    p = ptrs.cpu().numpy().view(np.uint64)
    assert np.all(p != 0)
    g = gpu.global_allocator()
    H.check_ptrs(p, np.full(len(p), 64, np.uint64), g.pool_base, g.num_segments * S)
    gpu.global_free_bulk(ptrs)
    torch.cuda.synchronize()
    gpu.print_global_stats()
    g.drain()
    check_fresh(g)
    gpu.free_global_allocator()

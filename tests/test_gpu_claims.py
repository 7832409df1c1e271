"""Claim-level parity of the batched forms with the oracle (GPU).

A workgroup batch (batch_malloc, batched.hpp) and a slice-cache refill
(cache_malloc, slice_cache.hpp) run on a deterministically booted instance
(one partition, slot 0, smallest-id segments).  Which request gets which
slice depends on scheduling, but the set of slices handed out per size
class must be exactly the oracle's coalesced claims of that many
consecutive slices of the class's pinned block (gallatin.cuh:865-1078,
block.cuh:245-281), including a claim that runs past slice 4095 and
replaces the block mid-batch.  An off-by-one in the batch's range
arithmetic (malloc_take) or the cache's (cache_malloc) changes a set.
Afterwards every allocation is freed and the instance drains to fresh.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
from tests import helpers as H

pytestmark = pytest.mark.gpu

G = H.golden()
LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "claims", "libclaim_kernels.so")
QUICK = (6, 6, 6, 6)
POOL = 1 << 30


def _lib():
    if not os.path.exists(LIB):
        pytest.fail("tests/claims/libclaim_kernels.so is not built (run build())")
    L = ctypes.CDLL(LIB)
    L.claims_batch.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_uint32, ctypes.c_void_p,
                                                       ctypes.c_uint32]
    L.claims_cache.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_uint32, ctypes.c_void_p]
    return L


def _instances(gpu, pre16=0):
    """A deterministic GPU instance and the oracle, both with the first
    `pre16` 16-byte slices of class 0 taken (waves of up to 64); returns
    them and the offsets of those slices."""
    a = gpu.Gallatin(16 << 20, 16, 4096).generate_on_device(
        POOL, 42, print_info=False, flags=gpu.GAL_FLAG_DETERMINISTIC, limits=QUICK)
    o = oracle.Allocator(POOL, limits=QUICK)
    pre = []
    if pre16:
        nsteps = (pre16 + 63) // 64
        steps = np.zeros(nsteps, H.STEP_DTYPE)
        for k in range(nsteps):
            lanes = min(64, pre16 - 64 * k)
            steps[k]["mask"] = (1 << lanes) - 1
            steps[k]["arg"][:lanes] = 16
        got = a.run_wave_stream(steps)
        assert np.array_equal(got, o.run_stream(steps))
        pre = [int(got[64 * k + j]) for k in range(nsteps) for j in range(min(64, pre16 - 64 * k))]
        assert sorted(pre) == [16 * i for i in range(pre16)]
    return a, o, pre


def _oracle_claims(o, sizes):
    """The oracle's coalesced claims: per class in increasing order, waves
    of up to 64 requests of that class.  Returns offsets per class."""
    out = {}
    by_class = {}
    for s in sizes:
        by_class.setdefault(H.clog2(max(s, 16)) - 4, []).append(s)
    for t in sorted(by_class):
        reqs = by_class[t]
        got = []
        for k in range(0, len(reqs), 64):
            part = reqs[k:k + 64]
            r = o.malloc_wave((1 << len(part)) - 1, part)
            got.extend(int(x) for x in r[:len(part)])
        out[t] = sorted(got)
    return out


def _by_class(sizes, offs):
    out = {}
    for s, x in zip(sizes, offs):
        out.setdefault(H.clog2(max(s, 16)) - 4, []).append(int(x))
    return {t: sorted(v) for t, v in out.items()}


def _run(gpu, a, sizes, kind, threads=256):
    import torch
    sz = torch.tensor(sizes, dtype=torch.int64, device="cuda")
    out = torch.zeros(len(sizes), dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    if kind == "batch":
        rc = _lib().claims_batch(a.device_state(), sz.data_ptr(), len(sizes), out.data_ptr(),
                                 threads)
    else:
        rc = _lib().claims_cache(a.device_state(), sz.data_ptr(), len(sizes), out.data_ptr())
    assert rc == 0
    return out.cpu().numpy().view(np.uint64)


def _free_all_and_check_fresh(gpu, a, offs):
    import torch
    base = a.info()["pool_base"]
    ptrs = torch.tensor([base + int(x) for x in offs], dtype=torch.int64, device="cuda")
    a.free_bulk(ptrs)
    torch.cuda.synchronize()
    a.drain()
    H.check_fresh(a)


def _runs(runs):
    return sorted(r["first"] + r["stride"] * i for r in runs for i in range(r["count"]))


@pytest.mark.parametrize("kind,threads", [("batch", 256), ("batch", 1024), ("cache", 64)])
def test_mixed_classes_match_oracle_claims(gpu, kind, threads):
    rng = np.random.default_rng(7)
    n = threads
    pool = [16, 100, 128, 1000, 4000, 300, 2048]
    sizes = [int(x) for x in rng.choice(pool, n)]
    a, o, _ = _instances(gpu)
    offs = _run(gpu, a, sizes, kind, threads)
    assert np.all(offs != np.uint64(~np.uint64(0)))
    assert _by_class(sizes, offs) == _oracle_claims(o, sizes)
    _free_all_and_check_fresh(gpu, a, offs)
    a.free_on_device()


@pytest.mark.parametrize("name,kind", [("batch_claim_crossing_4095_1GiB", "batch"),
                                       ("cache_refill_crossing_4095_1GiB", "cache")])
def test_claim_crossing_slice_4095(gpu, name, kind):
    g = G[name]
    a, o, pre = _instances(gpu, g["pre_slices_class0"])
    sizes = [g["size"]] * g["requests"]
    offs = _run(gpu, a, sizes, kind, max(64, g["requests"]))
    got = sorted(int(x) for x in offs)
    assert got == _runs(g["runs"]), "hand-derived slices"
    assert got == _oracle_claims(o, sizes)[0], "oracle's coalesced claims"
    _free_all_and_check_fresh(gpu, a, list(offs) + pre)
    a.free_on_device()


def test_batch_runs_of_largest_slices(gpu):
    """Requests in (4 KiB, 32 KiB] are runs of 2-8 consecutive 4 KiB slices
    of the largest class (gallatin.cuh:1210-1217).  A workgroup batch claims
    them with the class's single slices; a run that would straddle two of
    the claim's ranges (two blocks) credits its part back and takes the
    per-wave path.  Two batches of 1,024 requests, about 4,600 slices each,
    so claims cross block ends: every run lies in one block (a 16 MiB
    segment here), no two allocations overlap, singles are slice-aligned,
    and after all frees the instance drains to fresh, which fails if any
    pre-credit or hand-back is off by one."""
    rng = np.random.default_rng(31)
    a = gpu.Gallatin(16 << 20, 16, 4096).generate_on_device(POOL, 7, print_info=False)
    spans, offs = [], []
    for _ in range(2):
        sizes = [int(x) for x in np.where(rng.random(1024) < 0.3, 4000,
                                          rng.integers(4097, 32769, 1024))]
        got = _run(gpu, a, sizes, "batch", 1024)
        assert np.all(got != np.uint64(~np.uint64(0))), "a request failed"
        for sz, off in zip(sizes, got):
            off = int(off)
            n = 1 if sz <= 4096 else 1 << (H.clog2(sz) - 12)
            assert off % 4096 == 0
            assert off >> 24 == (off + n * 4096 - 1) >> 24, "run across two blocks"
            spans.append((off, off + n * 4096))
            offs.append(off)
    spans.sort()
    assert all(spans[i][1] <= spans[i + 1][0] for i in range(len(spans) - 1)), "overlap"
    assert a.info()["errors"] == 0
    _free_all_and_check_fresh(gpu, a, offs)
    a.free_on_device()

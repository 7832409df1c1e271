"""Batched request streams on the GPU: the workgroup-batched path, the LDS
slice-cache path (batched=2) and the plain path serve the same Zipf stream,
eagerly and as replays of one captured hipGraph; every replay issues every
malloc, frees every successful one exactly once with its tag intact, leaves
the pointer table empty, and the drained instance equals the fresh state
(for the cache: every cached slice left unused went back to its block)."""
import numpy as np
import pytest

from gallatin_amd import batch_streams as B
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_stream(gpu):
    ops, off, nslots = B.zipf_stream(11, nbatches=48, max_count=1 << 13)
    B.check_stream(ops, off, nslots)
    return gpu.BatchStream(ops, off, nslots)


def _check(gpu, c, stream, reps=1, misses_ok=False):
    assert c[gpu.CTR_MALLOCS] == reps * stream.n_mallocs
    assert c[gpu.CTR_BAD] == 0
    assert c[gpu.CTR_TIMEOUTS] == 0
    assert c[gpu.CTR_PAIRS] == c[gpu.CTR_MALLOCS] - c[gpu.CTR_MISSES]
    if not misses_ok:
        assert c[gpu.CTR_MISSES] == 0
        # every free found its pointer: the table writes of earlier batches
        # were visible to it
        assert c[gpu.CTR_FREES_SKIPPED] == 0


@pytest.mark.parametrize("seg,mn,batched,threads", [
    (16 << 20, 16, True, 1024), (16 << 20, 16, True, 256), (16 << 20, 16, False, 256),
    (4 << 20, 128, True, 512), (64 << 20, 16, True, 1024),
    (16 << 20, 16, 2, 1024), (16 << 20, 16, 2, 256), (4 << 20, 128, 2, 512),
    (64 << 20, 256, True, 1024), (64 << 20, 256, 2, 1024), (64 << 20, 256, False, 256)])
def test_eager_stream(gpu, small_stream, seg, mn, batched, threads):
    import torch
    a = gpu.Gallatin(seg, mn, min(seg // 4096, 4096)).generate_on_device(
        8 << 30, 5, print_info=False)
    ctr = gpu.run_batches(a, small_stream, threads=threads, batched=batched)
    torch.cuda.synchronize()
    _check(gpu, ctr.cpu().tolist(), small_stream)
    assert not torch.any(small_stream.ptrs != 0), "pointer table not empty"
    a.drain()
    H.check_fresh(a)
    a.free_on_device()


@pytest.mark.parametrize("seg,mn", [(16 << 20, 16), (64 << 20, 256)])
@pytest.mark.parametrize("batched", [True, False, 2])
def test_graph_replays(gpu, small_stream, batched, seg, mn):
    """Three replays of the captured graph, each followed by eager work on
    the same instance (an eager run of the stream) and a synchronisation;
    every replay is checked on its own."""
    import torch
    a = gpu.Gallatin(seg, mn, 4096).generate_on_device(8 << 30, 6, print_info=False)
    ctr = torch.zeros(gpu.CTR_COUNT, dtype=torch.int64, device="cuda")
    plan = gpu.BatchPlan(a, small_stream, ctr, batched=batched)
    for rep in range(3):
        ctr.zero_()
        torch.cuda.synchronize()
        plan.launch()
        torch.cuda.synchronize()
        _check(gpu, ctr.cpu().tolist(), small_stream)
        assert not torch.any(small_stream.ptrs != 0), f"replay {rep}: table not empty"
        eager = gpu.run_batches(a, small_stream, batched=not batched)
        torch.cuda.synchronize()
        _check(gpu, eager.cpu().tolist(), small_stream)
    plan.destroy()
    a.drain()
    H.check_fresh(a)
    a.free_on_device()


def test_autotune_sweeps_both_batched_forms(gpu, small_stream):
    """One geometry and workgroup shape: the sweep measures workgroup
    batches, the LDS slice cache and the per-thread path, every point
    without violations, and names a best batched point."""
    from gallatin_amd import autotune
    best, table = autotune.tune(small_stream, 8 << 30, seed=3, reps=1, warmup=0,
                                segments=(16 << 20,), smallest=(16,), threads=(256,))
    assert {r["form"] for r in table} == {"workgroup", "cache", "per-thread"}
    assert all(r["violations"] == 0 for r in table), table
    assert best is not None and best["path"] == "batched"


def test_per_batch_launch_measure(gpu, small_stream):
    """The bench's per-batch-launch timing (one launch per batch and one for
    the cleanup) counts every pair of the stream, for the per-thread and the
    workgroup form, on one instance; the instance drains to the fresh
    state."""
    from gallatin_amd import autotune
    a = gpu.Gallatin(16 << 20, 16, 4096).generate_on_device(8 << 30, 9, print_info=False)
    for batched, threads in ((False, 256), (1, 1024)):
        r = autotune.measure_launches(a, small_stream, batched, threads, reps=2, warmup=1)
        assert r["violations"] == 0 and r["misses"] == 0, r
        assert r["pairs_per_s"] > 0 and r["p50_step_ms"] <= r["p99_step_ms"]
    a.drain()
    H.check_fresh(a)
    a.free_on_device()


@pytest.mark.parametrize("batched", [True, False])
def test_timed_plan_splits_wave_time(gpu, small_stream, batched):
    """A plan made with timed=True (GAL_BATCH_TIMED) serves the stream the
    same way and adds the stream kernel's wave time split: malloc rounds,
    waits and frees are parts of the waves' lifetime.  An untimed plan
    leaves those counters at zero."""
    import torch
    a = gpu.Gallatin(16 << 20, 16, 4096).generate_on_device(8 << 30, 8, print_info=False)
    for timed in (False, True):
        ctr = torch.zeros(gpu.CTR_COUNT, dtype=torch.int64, device="cuda")
        plan = gpu.BatchPlan(a, small_stream, ctr, batched=batched, timed=timed)
        plan.launch()
        torch.cuda.synchronize()
        plan.destroy()
        c = ctr.cpu().tolist()
        _check(gpu, c, small_stream)
        parts = [c[gpu.CTR_MALLOC_TICKS], c[gpu.CTR_WAIT_TICKS], c[gpu.CTR_FREE_TICKS]]
        if not timed:
            assert parts == [0, 0, 0] and c[gpu.CTR_WAVE_TICKS] == 0
        else:
            assert c[gpu.CTR_WAVE_TICKS] > 0 and parts[0] > 0
            assert sum(parts) <= c[gpu.CTR_WAVE_TICKS]
    a.drain()
    H.check_fresh(a)
    a.free_on_device()

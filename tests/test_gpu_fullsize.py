"""The benchmark configurations at their full stated size on the GPU.

Mixed-size churn: 8 GiB pool, 2^24 threads x 10 rounds of log-uniform sizes
in [1, 4096) with a 50 % replacement per round, every allocation filled with
its (thread, round) pattern and verified before it is freed, on the plain
per-thread path and the workgroup-batched path.  The pair count equals the
count the stream generator fixes, no pattern is broken, nothing misses, and
the drained instance equals the fresh state.

Batched persistent form: the benchmark's Zipf stream (256 batches, Zipf(1.2)
counts on [1, 2^20], 60 % malloc of lognormal(6, 2) bytes clipped to
[1, 16 MiB] / 40 % free of a uniformly chosen live pointer) on a 64 GiB pool,
replayed as one captured hipGraph in each form: every malloc issued, every
successful one freed exactly once with its tag intact, the pointer table
empty, fresh after a drain.

(The serving configuration at full size is in test_gpu_serving.py, the plain
small and the block + multi-segment configurations in test_gpu_allocator.py.)
"""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

SEED = 1234


@pytest.fixture(scope="module")
def churn_pairs():
    return H.expected_churn_pairs(SEED, 1 << 24, 10)


@pytest.mark.parametrize("path", ["plain", "batched"])
def test_mixed_churn_full_size(gpu, churn_pairs, path):
    import torch
    a = gpu.Gallatin(16 << 20, 16, 4096).generate_on_device(8 << 30, 42, print_info=False,
                                                           slots_min=32)
    if path == "batched":
        ctr = a.run_churn_batched(SEED, 1 << 24, 10, threads=1024)
    else:
        ctr = a.run_churn(SEED, 1 << 24, 10)
    torch.cuda.synchronize()
    c = ctr.cpu().tolist()
    assert c[gpu.CTR_BAD] == 0
    assert c[gpu.CTR_MISSES] == 0
    assert c[gpu.CTR_PAIRS] == churn_pairs
    assert a.info()["errors"] == 0
    a.drain()
    H.check_fresh(a)
    a.free_on_device()


@pytest.fixture(scope="module")
def zipf_full(gpu):
    from gallatin_amd import batch_streams as B
    ops, off, nslots = B.zipf_stream(SEED)
    assert len(off) == 257
    gpu.check_batch_stream(ops, off, nslots, persistent=True)
    return gpu.BatchStream(ops, off, nslots)


@pytest.mark.parametrize("batched,threads", [(1, 1024), (2, 1024), (0, 256)])
def test_batched_zipf_full_size(gpu, zipf_full, batched, threads):
    import torch
    a = gpu.Gallatin(64 << 20, 128, 4096).generate_on_device(
        64 << 30, SEED, print_info=False, slots_min=32, flags=gpu.GAL_FLAG_NO_PAYLOAD_FENCE)
    ctr = torch.zeros(gpu.CTR_COUNT, dtype=torch.int64, device="cuda")
    plan = gpu.BatchPlan(a, zipf_full, ctr, threads=threads, batched=batched)
    for rep in range(2):
        ctr.zero_()
        torch.cuda.synchronize()
        plan.launch()
        torch.cuda.synchronize()
        c = ctr.cpu().tolist()
        assert c[gpu.CTR_MALLOCS] == zipf_full.n_mallocs, rep
        assert c[gpu.CTR_BAD] == 0 and c[gpu.CTR_TIMEOUTS] == 0, rep
        assert c[gpu.CTR_MISSES] == 0 and c[gpu.CTR_FREES_SKIPPED] == 0, rep
        assert c[gpu.CTR_PAIRS] == zipf_full.n_mallocs, rep
        assert not torch.any(zipf_full.ptrs != 0), f"replay {rep}: table not empty"
    plan.destroy()
    assert a.info()["errors"] == 0
    a.drain()
    H.check_fresh(a)
    a.free_on_device()


@pytest.mark.parametrize("batched", [1, 0])
def test_batched_zipf_small_segments(gpu, zipf_full, batched):
    """The full stream at 4 MiB segments and 256 B smallest slices, a point
    of the autotuner's grid: the largest class has 1 KiB slices and one
    block per segment, and a third of the requests are runs of up to 512 of
    them, so a workgroup's claims ask for many blocks' worth while blocks
    fill and are replaced all the time.  Claims that added their whole
    demand to a full block once carried its count into the tag here
    (allocator errors).  No allocator error, no bad tag, every successful
    malloc freed once, and fresh after the drain."""
    import torch
    a = gpu.Gallatin(4 << 20, 256, 1024).generate_on_device(
        64 << 30, SEED, print_info=False, slots_min=32, flags=gpu.GAL_FLAG_NO_PAYLOAD_FENCE)
    ctr = torch.zeros(gpu.CTR_COUNT, dtype=torch.int64, device="cuda")
    plan = gpu.BatchPlan(a, zipf_full, ctr, threads=1024 if batched else 256, batched=batched)
    for rep in range(3):
        ctr.zero_()
        torch.cuda.synchronize()
        plan.launch()
        torch.cuda.synchronize()
        c = ctr.cpu().tolist()
        assert c[gpu.CTR_MALLOCS] == zipf_full.n_mallocs, rep
        assert c[gpu.CTR_BAD] == 0 and c[gpu.CTR_TIMEOUTS] == 0, rep
        assert c[gpu.CTR_PAIRS] == c[gpu.CTR_MALLOCS] - c[gpu.CTR_MISSES], rep
        assert not torch.any(zipf_full.ptrs != 0), f"replay {rep}: table not empty"
        assert a.info()["errors"] == 0, (rep, a.stats())
    plan.destroy()
    a.drain()
    H.check_fresh(a)
    a.free_on_device()

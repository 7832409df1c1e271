"""The entry point that allocates nothing (gal_run_stream_static) against
the persistent plan it varies (gal_batch_plan_create + launch): on the same
Zipf stream its event counters equal the plan's exactly (every malloc
issued, every successful one freed once with its tag intact, no wait ran
out; the wall-clock counters and wait polls differ from run to run), the
pointer table ends empty and the drained instance is fresh.  Two calls in a
row, the second on a shorter stream, so the shared workspace's leftovers
from the first call must not leak into the second.  Inputs beyond the
workspace are refused before any device work."""
import ctypes

import numpy as np
import pytest

from gallatin_amd import batch_streams as B
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _plan_counters(gpu, a, stream, batched, threads):
    import torch
    ctr = torch.zeros(gpu.CTR_COUNT, dtype=torch.int64, device="cuda")
    plan = gpu.BatchPlan(a, stream, ctr, threads=threads, batched=batched)
    plan.launch()
    torch.cuda.synchronize()
    plan.destroy()
    return ctr.cpu().tolist()


def _static_counters(gpu, a, stream, batched, threads):
    import torch
    ctr = torch.zeros(gpu.CTR_COUNT, dtype=torch.int64, device="cuda")
    gpu.run_stream_static(a, stream, ctr, threads=threads, batched=batched)
    torch.cuda.synchronize()
    return ctr.cpu().tolist()


@pytest.mark.parametrize("batched,threads", [(1, 1024), (0, 256), (2, 512)])
def test_static_matches_plan_twice(gpu, batched, threads):
    import torch
    a = gpu.Gallatin(16 << 20, 16, 4096).generate_on_device(8 << 30, 3, print_info=False)
    big = B.zipf_stream(21, nbatches=64, max_count=1 << 14)
    small = B.zipf_stream(22, nbatches=16, max_count=1 << 10)
    assert big[1][-1] > small[1][-1]
    for ops, off, nslots in (big, small):
        s = gpu.BatchStream(ops, off, nslots)
        exp = _plan_counters(gpu, a, s, batched, threads)
        got = _static_counters(gpu, a, s, batched, threads)
        assert got[:gpu.CTR_WAIT_TICKS] == exp[:gpu.CTR_WAIT_TICKS]
        assert got[gpu.CTR_MALLOCS] == s.n_mallocs and got[gpu.CTR_PAIRS] == s.n_mallocs
        assert got[gpu.CTR_BAD] == 0 and got[gpu.CTR_TIMEOUTS] == 0
        assert not torch.any(s.ptrs != 0), "pointer table not empty"
    assert a.info()["errors"] == 0
    a.drain()
    H.check_fresh(a)
    a.free_on_device()


def test_static_refuses_oversize_streams(gpu):
    import torch
    a = gpu.Gallatin(16 << 20, 16, 4096).generate_on_device(1 << 30, 3, print_info=False)
    lib = gpu.lib()
    ctr = torch.zeros(gpu.CTR_COUNT, dtype=torch.int64, device="cuda")
    ptrs = torch.zeros(4, dtype=torch.int64, device="cuda")
    # too many batches (all empty)
    off = np.zeros(70000, dtype=np.uint64)
    rc = lib.gal_run_stream_static(a.handle, ptrs.data_ptr(), off.ctypes.data, len(off) - 1,
                                   ptrs.data_ptr(), 4, ctr.data_ptr(), 256, 0, 0, None)
    assert rc == -1 and b"65535" in lib.gal_last_error()
    # one batch of more chunks than the workspace holds (1,024 ops each)
    off = np.array([0, 1024 * 131072 + 1], dtype=np.uint64)
    rc = lib.gal_run_stream_static(a.handle, ptrs.data_ptr(), off.ctypes.data, 1,
                                   ptrs.data_ptr(), 4, ctr.data_ptr(), 256, 0, 0, None)
    assert rc == -1 and b"131072" in lib.gal_last_error()
    torch.cuda.synchronize()
    assert int(ctr.abs().sum()) == 0
    a.free_on_device()

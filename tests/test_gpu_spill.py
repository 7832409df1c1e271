"""Pools outside HBM, spill to the host, and poison (gallatin.cuh:427-630,
global_allocator.cuh:100-357).

- A pool in pinned host memory costs no device memory beyond the metadata,
  and the plain and churn workloads run on it and on a managed pool with no
  misses and every pattern intact; afterwards each drains to its fresh
  state.
- Fused malloc takes the device pool first and the host pool once the
  device pool has no room: every pointer lies in exactly one pool, live
  allocations never overlap, and fused free returns each to its owner.
- Combined malloc places by the caller's flag.
- Poisoned allocations check clean; one byte written past the end of an
  allocation makes its check fail and no other; poisoned frees return
  every slot.

The reference's tests for these (global_exceed_test.cu:117-150,
poison_test.cu:97-194) print and always pass, so these invariants are
what pins the behaviour.  Its host-instance tests (global_host_test.cu and
global_combined_test.cu, 34-209: one-size mallocs through the global host
instance, freed through it) go through the same host instance that the
combined and fused tests here place allocations in and free them from.
"""
import ctypes

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu


def _plain_and_churn(gpu, a):
    import torch
    ctr = torch.zeros(gpu.CTR_COUNT, dtype=torch.int64, device="cuda")
    a.run_plain(1 << 18, 16, counters=ctr)
    torch.cuda.synchronize()
    c = ctr.cpu().numpy()
    assert c[gpu.CTR_PAIRS] == 1 << 18 and c[gpu.CTR_MISSES] == 0 and c[gpu.CTR_BAD] == 0
    ctr.zero_()
    n, rounds, seed = 1 << 16, 4, 5
    a.run_churn(seed, n, rounds, counters=ctr)
    torch.cuda.synchronize()
    c = ctr.cpu().numpy()
    assert c[gpu.CTR_BAD] == 0 and c[gpu.CTR_MISSES] == 0
    assert c[gpu.CTR_PAIRS] == H.expected_churn_pairs(seed, n, rounds)
    a.drain()
    H.check_fresh(a)


def test_host_pool(gpu):
    import torch
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    a = gpu.Gallatin(16 << 20, 16, 4096).generate_on_device_host(4 << 30, 42, print_info=False)
    try:
        free1, _ = torch.cuda.mem_get_info()
        assert free0 - free1 < (1 << 30), "a 4 GiB host pool took device memory"
        _plain_and_churn(gpu, a)
    finally:
        a.free_on_device()


def test_managed_pool(gpu):
    """Runs where the device reports managed-memory support; elsewhere the
    allocator refuses the pool (checked here) and the workload is skipped."""
    try:
        a = gpu.Gallatin(16 << 20, 16, 4096).generate_on_device_managed(1 << 30, 43,
                                                                        print_info=False)
    except gpu.GallatinError as e:
        assert "not supported" in str(e)
        pytest.skip("the device reports no managed-memory support")
    try:
        _plain_and_churn(gpu, a)
    finally:
        a.free_on_device()


def test_host_and_managed_together_refused(gpu):
    with pytest.raises(gpu.GallatinError):
        gpu.Gallatin().generate_on_device(1 << 30, 1, print_info=False,
                                          flags=gpu.GAL_FLAG_POOL_HOST
                                          | gpu.GAL_FLAG_POOL_MANAGED)


@pytest.fixture
def combined(gpu):
    gpu.init_global_allocator_combined(512 << 20, 1 << 30, 11, print_info=False)
    yield gpu
    for g in (gpu.global_allocator(), gpu.global_host_allocator()):
        g.drain()
        H.check_fresh(g)
    gpu.free_global_allocator_combined()


def _split(ptrs, a):
    lo, hi = a.pool_base, a.pool_base + a.num_segments * (16 << 20)
    return (ptrs >= np.uint64(lo)) & (ptrs < np.uint64(hi))


def test_fused_spills_to_host(combined):
    import torch
    g, h = combined.global_allocator(), combined.global_host_allocator()
    n, size = 1 << 17, 4096  # 512 MiB: more than the 512 MiB device pool can hold
    sizes = np.full(n, size, dtype=np.uint64)
    ptrs_d = combined.global_malloc_fused_bulk(torch.full((n,), size, dtype=torch.int64,
                                                          device="cuda"))
    torch.cuda.synchronize()
    p = ptrs_d.cpu().numpy().view(np.uint64)
    assert np.all(p != 0)
    on_dev, on_host = _split(p, g), _split(p, h)
    assert np.all(on_dev ^ on_host), "a pointer in neither or both pools"
    assert on_dev.sum() > 0 and on_host.sum() > 0
    H.check_ptrs(p[on_dev], sizes[on_dev], g.pool_base, g.num_segments * (16 << 20))
    H.check_ptrs(p[on_host], sizes[on_host], h.pool_base, h.num_segments * (16 << 20))
    combined.global_free_fused_bulk(ptrs_d)
    torch.cuda.synchronize()


def test_combined_places_by_flag(combined):
    import torch
    g, h = combined.global_allocator(), combined.global_host_allocator()
    sizes = torch.randint(1, 4097, (1 << 14,), dtype=torch.int64, device="cuda")
    pd = combined.global_malloc_combined_bulk(sizes, on_host=False)
    ph = combined.global_malloc_combined_bulk(sizes, on_host=True)
    torch.cuda.synchronize()
    a, b = pd.cpu().numpy().view(np.uint64), ph.cpu().numpy().view(np.uint64)
    assert np.all(a != 0) and np.all(b != 0)
    assert np.all(_split(a, g)) and np.all(_split(b, h))
    combined.global_free_combined_bulk(pd, on_host=False)
    combined.global_free_combined_bulk(ph, on_host=True)
    torch.cuda.synchronize()


def test_poison(gpu):
    import torch
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    gpu.init_global_allocator(1 << 30, 13, print_info=False)
    try:
        rng = np.random.default_rng(3)
        n = 4096
        sizes = rng.integers(1, 3000, n)
        ptrs = gpu.global_malloc_poison_bulk(torch.from_numpy(sizes).cuda())
        ok = gpu.global_check_poison_bulk(ptrs)
        torch.cuda.synchronize()
        p = ptrs.cpu().numpy().view(np.uint64)
        assert np.all(p != 0) and bool(ok.all())
        # the byte just past the end of 16 allocations is flipped
        bad = rng.choice(n, 16, replace=False)
        byte = ctypes.c_uint8()
        torch.cuda.synchronize()
        for i in bad:
            at = int(p[i]) + int(sizes[i])
            assert hip.hipMemcpy(ctypes.addressof(byte), at, 1, 2) == 0  # device to host
            byte.value ^= 0xFF
            assert hip.hipMemcpy(at, ctypes.addressof(byte), 1, 1) == 0  # host to device
        ok = gpu.global_check_poison_bulk(ptrs).cpu().numpy()
        exp = np.ones(n, dtype=np.uint8)
        exp[bad] = 0
        assert np.array_equal(ok, exp)
        gpu.global_free_poison_bulk(ptrs)
        torch.cuda.synchronize()
        g = gpu.global_allocator()
        g.drain()
        H.check_fresh(g)
    finally:
        gpu.free_global_allocator()

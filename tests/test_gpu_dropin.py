"""The C++ headers as a drop-in: a user translation unit written like the
reference's own tests (tests/src/global_test.cu:33-120) includes
global_allocator.hpp, defines its own global allocator symbol, calls
global_malloc / global_free and Gallatin<>::malloc / free from its kernels,
and links libgallatin_hip.so for the host side."""
import ctypes
import os

import pytest

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dropin", "libuser_kernels.so")


def _lib():
    if not os.path.exists(LIB):
        pytest.fail("tests/dropin/libuser_kernels.so is not built (run build())")
    L = ctypes.CDLL(LIB)
    L.user_run.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                           ctypes.POINTER(ctypes.c_ulonglong)]
    L.user_run.restype = ctypes.c_int
    L.user_structures.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_ulonglong)]
    L.user_structures.restype = ctypes.c_int
    return L


def test_user_library_loads_and_exports():
    L = _lib()
    for name in ("user_run", "user_structures", "user_table_accessors", "user_churn",
                 "user_single_thread", "user_host_pool"):
        assert hasattr(L, name), name


@pytest.mark.gpu
@pytest.mark.parametrize("size,n", [(16, 1 << 18), (100, 1 << 18), (4096, 1 << 16)])
def test_user_kernels_global_and_own_instance(size, n):
    """n live allocations of `size` fit a 1 GiB pool with room to spare
    (2^16 x 4 KiB = 256 MiB), so every malloc must succeed."""
    import torch
    torch.cuda.set_device(0)
    counts = (ctypes.c_ulonglong * 4)()
    rc = _lib().user_run(1 << 30, n, size, counts)
    assert rc == 0
    assert list(counts) == [0, 0, 0, 0], "misses / errors in the user kernels"


@pytest.mark.gpu
def test_user_kernels_data_structures():
    """The user unit includes the data-structure headers and runs each the
    way the reference's tests do: a vector's items at unique positions, hash
    and chaining tables queried back after inserts (keys tid + 1, values
    tid), and two queues that hand back every item exactly once."""
    import torch
    torch.cuda.set_device(0)
    counts = (ctypes.c_ulonglong * 5)()
    rc = _lib().user_structures(1 << 16, counts)
    assert rc == 0
    assert list(counts) == [0] * 5, "vector, table, chain, queue, block queue errors"


@pytest.mark.gpu
@pytest.mark.parametrize("size", [16, 100, 4096])
def test_user_kernels_table_accessors(size):
    """The reference's mainline test body (tests/src/global_test_noprint.cu:
    81-185) compiles against the headers unchanged apart from the include
    line: it reaches segment, tree id, slice offset and block counters
    through global_gallatin->table and allocation_to_offset, and those agree
    with the geometry; ids read back and no block counts past 4096."""
    import torch
    torch.cuda.set_device(0)
    L = _lib()
    L.user_table_accessors.argtypes = [ctypes.c_uint64] * 3 + [ctypes.POINTER(ctypes.c_ulonglong)]
    counts = (ctypes.c_ulonglong * 4)()
    assert L.user_table_accessors(1 << 30, 1 << 16, size, counts) == 0
    assert list(counts) == [0, 0, 0, 0], "misses, accessor, id, block-count errors"


@pytest.mark.gpu
@pytest.mark.parametrize("num_bytes,threads,rounds,min_size,max_size", [
    (1000 * (16 << 20), 1 << 20, 10, 16, 4096),   # the test's own defaults
    (4 << 30, 1 << 18, 10, 16, 4096),             # one partition
    (8 << 30, 1 << 16, 4, 16, 65536),             # runs of largest slices too
])
def test_user_churn_kernel(num_bytes, threads, rounds, min_size, max_size):
    """The reference's churn test kernel (tests/src/gallatin_churn.cu:37-165)
    written against the headers in the user unit: each thread mallocs a
    hashed size every round, exchanges its id into the first word, which
    must hold 0, exchanges 0 back, which must return its id, and frees.
    Every allocation is freed in its round, so the pools leave ample room:
    no miss, no double malloc, no allocator error."""
    import torch
    torch.cuda.set_device(0)
    L = _lib()
    L.user_churn.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64,
                             ctypes.c_uint64, ctypes.POINTER(ctypes.c_ulonglong)]
    L.user_churn.restype = ctypes.c_int
    counts = (ctypes.c_ulonglong * 4)()
    assert L.user_churn(num_bytes, threads, rounds, min_size, max_size, counts) == 0
    assert list(counts) == [0, 0, 0, 0], "misses, double mallocs (claim, return), errors"


@pytest.mark.gpu
@pytest.mark.parametrize("size,n", [(16, 10000), (4096, 10000), (1 << 20, 1000),
                                    (32 << 20, 100)])
def test_user_single_thread(size, n):
    """The reference's single-thread test (tests/src/gallatin_single_thread.cu:
    38-140) in the user unit: one thread mallocs n allocations of one size in
    a row and then frees them, two rounds, through the global allocator.
    Slices (16 B, 4 KiB: several blocks, 4 KiB ones several segments), whole
    blocks (1 MiB) and two-segment runs (32 MiB): every index reads back, no
    miss, no error, and after a drain every segment is free."""
    import torch
    torch.cuda.set_device(0)
    L = _lib()
    L.user_single_thread.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64,
                                     ctypes.POINTER(ctypes.c_ulonglong)]
    L.user_single_thread.restype = ctypes.c_int
    counts = (ctypes.c_ulonglong * 4)()
    assert L.user_single_thread(n, 2, size, counts) == 0
    assert list(counts) == [0, 0, 0, 0], "misses, bad indices, errors, segments not free"


@pytest.mark.gpu
@pytest.mark.parametrize("size", [64, 4096])
def test_user_host_pool_overlaps(size):
    """The reference's host-pool test (tests/src/gallatin_host_test.cu:40-260)
    in the user unit: an instance from generate_on_device_host, 2^18
    allocations of one slice class written with their ids, and that test's
    overlap check, which counts allocations per block from
    alloc->table->memory: nothing missed, every id reads back, no block
    holds more than 4096 and the blocks hold every allocation."""
    import torch
    torch.cuda.set_device(0)
    L = _lib()
    L.user_host_pool.argtypes = [ctypes.c_uint64] * 3 + [ctypes.POINTER(ctypes.c_ulonglong)]
    L.user_host_pool.restype = ctypes.c_int
    n = 1 << 18
    counts = (ctypes.c_ulonglong * 5)()
    assert L.user_host_pool(4 << 30, n, size, counts) == 0
    assert list(counts) == [0, 0, 0, n, 0], "misses, ids, blocks, counted, errors"

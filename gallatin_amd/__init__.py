"""Python host interface of the MI355X Gallatin allocator (libgallatin_hip.so).

Mirrors the reference's host API (include/gallatin/allocators/gallatin.cuh
and global_allocator.cuh of saltsystemslab/gallatin) over the C ABI declared
in include/gallatin_capi.h:

    alloc = Gallatin(16 << 20, 16, 4096).generate_on_device(num_bytes, seed)
    alloc.print_info();  alloc.free_on_device()
    init_global_allocator(num_bytes, seed); print_global_stats(); free_global_allocator()

plus the workload kernels the reference's tests run (bulk malloc/free, the
fixed-size and churn tests) and a single-wave deterministic stream.  Device
buffers are torch tensors on ``cuda``; torch is only the memory/stream
plumbing.  Everything fails loudly: a missing library raises ImportError on
first use, a failing call raises GallatinError.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgallatin_hip.so")
FAIL = (1 << 64) - 1

GAL_FLAG_DETERMINISTIC = 1
GAL_FLAG_NO_PAYLOAD_FENCE = 2
GAL_FLAG_POOL_HOST = 4
GAL_FLAG_POOL_MANAGED = 8
GAL_FLAG_LAZY_BOOT = 16

CTR_PAIRS, CTR_MISSES, CTR_BAD, CTR_BYTES, CTR_MALLOCS = 0, 1, 2, 3, 4
CTR_TIMEOUTS, CTR_FREES_SKIPPED = 5, 6
CTR_WAIT_TICKS, CTR_WAVE_TICKS, CTR_WAIT_POLLS, CTR_MALLOC_TICKS, CTR_FREE_TICKS = 7, 8, 9, 10, 11
CTR_COUNT = 12
GAL_BATCH_TIMED = 0x100  # or-ed into `batched`: the persistent kernel times its waves

_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32
_p = ctypes.c_void_p

# symbols declared in include/gallatin_capi.h
CAPI_SYMBOLS = [
    "gal_config_default", "gal_create", "gal_destroy", "gal_device_state",
    "gal_from_device_state", "gal_print_info", "gal_get_info", "gal_drain",
    "gal_reboot", "gal_snapshot", "gal_tree_words", "gal_init_global",
    "gal_free_global", "gal_print_global_stats", "gal_global_handle",
    "gal_malloc_bulk", "gal_free_bulk", "gal_global_malloc_bulk",
    "gal_global_free_bulk", "gal_run_wave_stream", "gal_run_plain",
    "gal_run_churn", "gal_run_churn_batched", "gal_run_churn_word", "gal_run_big",
    "gal_bench_atomic",
    "gal_bench_copy",
    "gal_bench_scatter", "gal_bench_tree",
    "gal_last_error", "gal_get_stats", "gal_run_batches", "gal_batch_plan_create",
    "gal_batch_plan_launch", "gal_batch_plan_destroy", "gal_check_batch_stream",
    "gal_run_stream_static",
    "gal_vector_create", "gal_vector_destroy", "gal_vector_insert", "gal_vector_export",
    "gal_table_create", "gal_table_destroy", "gal_table_insert", "gal_table_query",
    "gal_table_remove", "gal_table_stats", "gal_serving_create", "gal_serving_step",
    "gal_serving_check", "gal_serving_table", "gal_serving_destroy", "gal_serving_gate",
    "gal_queue_create", "gal_queue_destroy", "gal_queue_enqueue", "gal_queue_dequeue",
    "gal_chain_create", "gal_chain_destroy", "gal_chain_insert", "gal_chain_query",
    "gal_chain_stats", "gal_chain_lengths", "gal_chain_print",
    "gal_bqueue_create", "gal_bqueue_destroy", "gal_bqueue_enqueue", "gal_bqueue_dequeue",
    "gal_bqueue_reclaim", "gal_bqueue_stats",
    "gal_init_global_host", "gal_free_global_host", "gal_print_global_stats_host",
    "gal_global_host_handle", "gal_global_malloc_fused_bulk", "gal_global_free_fused_bulk",
    "gal_global_malloc_combined_bulk", "gal_global_free_combined_bulk",
    "gal_global_malloc_poison_bulk", "gal_global_check_poison_bulk",
    "gal_global_free_poison_bulk",
]


class GallatinError(RuntimeError):
    pass


class Config(ctypes.Structure):
    _fields_ = [("max_bytes", _u64), ("bytes_per_segment", _u64),
                ("smallest", _u64), ("biggest", _u64), ("seed", _u64),
                ("flags", _u32), ("slots0", _u32), ("slots_min", _u32),
                ("print_info", _u32), ("request_attempts", _u32),
                ("slice_attempts", _u32), ("block_attempts", _u32),
                ("segment_attempts", _u32), ("partitions", _u32), ("pad", _u32)]


class Info(ctypes.Structure):
    _fields_ = [("num_segments", _u64), ("bytes_per_segment", _u64),
                ("smallest", _u64), ("biggest", _u64),
                ("blocks_per_segment", _u64), ("num_trees", _u32),
                ("total_slots", _u32), ("partitions", _u32), ("pad", _u32),
                ("segments_free", _u64),
                ("segments_formatted", _u64), ("segments_multi", _u64),
                ("tree_owned", _u64 * 24), ("errors", _u64),
                ("overhead_bytes", _u64), ("pool_base", _u64),
                ("class_segments", _u64 * 24), ("class_blocks_free", _u64 * 24),
                ("class_blocks_used", _u64 * 24)]


STEP_DTYPE = np.dtype([("kind", "<u4"), ("pad", "<u4"), ("mask", "<u8"),
                       ("arg", "<u8", (64,))])
# gal_batch_op
BATCH_OP_DTYPE = np.dtype([("arg", "<u8"), ("kind", "<u4"), ("slot", "<u4")])

_lib = None


def lib():
    """Load libgallatin_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m gallatin_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.gal_config_default.argtypes = [ctypes.POINTER(Config), _u64, _u64]
        L.gal_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(_p)]
        L.gal_destroy.argtypes = [_p]
        L.gal_device_state.restype = _p
        L.gal_device_state.argtypes = [_p]
        L.gal_from_device_state.restype = _p
        L.gal_from_device_state.argtypes = [_p]
        L.gal_print_info.argtypes = [_p]
        L.gal_get_info.argtypes = [_p, ctypes.POINTER(Info)]
        L.gal_get_stats.argtypes = [_p, _p, _u32]
        L.gal_drain.argtypes = [_p, _p]
        L.gal_reboot.argtypes = [_p, _p]
        L.gal_snapshot.argtypes = [_p, _p, _p, _p]
        L.gal_tree_words.restype = ctypes.c_int64
        L.gal_tree_words.argtypes = [_p, ctypes.c_int, ctypes.c_int, _p, _u64, _p, _p]
        L.gal_init_global.argtypes = [_u64, _u64, ctypes.c_int]
        L.gal_global_handle.restype = _p
        L.gal_malloc_bulk.argtypes = [_p, _p, _u64, _p, _u32, _p]
        L.gal_free_bulk.argtypes = [_p, _p, _u64, _u32, _p]
        L.gal_global_malloc_bulk.argtypes = [_p, _u64, _p, _u32, _p]
        L.gal_global_free_bulk.argtypes = [_p, _u64, _u32, _p]
        L.gal_run_wave_stream.argtypes = [_p, _p, _u64, _p, _p]
        L.gal_run_plain.argtypes = [_p, _u64, _u64, _p, _u32, _p]
        L.gal_run_churn.argtypes = [_p, _u64, _u64, _u32, _p, _u32, _p]
        L.gal_run_churn_batched.argtypes = [_p, _u64, _u64, _u32, _p, _u32, _u32, _p]
        L.gal_run_churn_word.argtypes = [_p, _u64, _u64, _u32, _p, _u32, _u32, _u32, _p]
        L.gal_run_big.argtypes = [_p, _u64, _u32, _u64, _u64, _p, _p, _p]
        L.gal_bench_atomic.argtypes = [_p, _u64, _u64, _u32, _u32, _p]
        L.gal_bench_copy.argtypes = [_p, _p, _u64, _p]
        L.gal_bench_scatter.argtypes = [_p, _p, _u64, _u64, _p]
        L.gal_bench_tree.argtypes = [_p, _u32, _u32, _u32, _p, _p]
        L.gal_run_batches.argtypes = [_p, _p, _p, _u32, _p, _u64, _p, _u32, _u32, _u32, _p]
        L.gal_batch_plan_create.argtypes = [_p, _p, _p, _u32, _p, _u64, _p, _u32, _u32,
                                            _u32, ctypes.POINTER(_p)]
        L.gal_batch_plan_launch.argtypes = [_p, _p]
        L.gal_check_batch_stream.argtypes = [_p, _p, _u32, _u64, ctypes.c_int]
        L.gal_run_stream_static.argtypes = [_p, _p, _p, _u32, _p, _u64, _p, _u32, _u32, _u32,
                                            _p]
        L.gal_batch_plan_destroy.argtypes = [_p]
        L.gal_vector_create.argtypes = [ctypes.POINTER(_p)]
        L.gal_vector_destroy.argtypes = [_p]
        L.gal_vector_insert.argtypes = [_p, _p, _u64, _p, ctypes.c_int, _p]
        L.gal_vector_export.restype = ctypes.c_int64
        L.gal_vector_export.argtypes = [_p, _p, _u64]
        L.gal_table_create.argtypes = [ctypes.c_int, _u64, ctypes.POINTER(_p)]
        L.gal_table_destroy.argtypes = [_p]
        L.gal_table_insert.argtypes = [_p, _p, _p, _u64, _p, _p]
        L.gal_table_query.argtypes = [_p, _p, _u64, _p, _p, _p]
        L.gal_table_remove.argtypes = [_p, _p, _u64, _p, _p]
        L.gal_table_stats.argtypes = [_p, _p, _p, _p, _p]
        L.gal_serving_create.argtypes = [_u32, _u64, _p, ctypes.POINTER(_p)]
        L.gal_serving_step.argtypes = [_p, _u32, _u32, _u64, _u32, _u32, _u32, _p, _p]
        L.gal_serving_gate.argtypes = [_p, _p, _p]
        L.gal_serving_check.argtypes = [_p, _p, _p, _p, _p]
        L.gal_serving_table.argtypes = [_p, _p, _p, _p]
        L.gal_serving_destroy.argtypes = [_p]
        L.gal_queue_create.argtypes = [_p, ctypes.c_int, ctypes.POINTER(_p)]
        L.gal_queue_destroy.argtypes = [_p]
        L.gal_queue_enqueue.argtypes = [_p, _p, _u64, _p, ctypes.c_int, _p]
        L.gal_queue_dequeue.argtypes = [_p, _p, _p, _u64, ctypes.c_int, _p]
        L.gal_chain_create.argtypes = [ctypes.c_int, _u64, _u64, _u64, ctypes.c_int,
                                       ctypes.POINTER(_p)]
        L.gal_chain_destroy.argtypes = [_p]
        L.gal_chain_insert.argtypes = [_p, _p, _p, _u64, _p, _p]
        L.gal_chain_query.argtypes = [_p, _p, _u64, _p, _p, _p]
        L.gal_chain_stats.argtypes = [_p, _p, _p, _p, _p]
        L.gal_chain_lengths.argtypes = [_p, _p, _p]
        L.gal_chain_print.argtypes = [_p]
        L.gal_bqueue_create.argtypes = [ctypes.c_int, ctypes.POINTER(_p)]
        L.gal_bqueue_destroy.argtypes = [_p]
        L.gal_bqueue_enqueue.argtypes = [_p, _p, _u64, _p, ctypes.c_int, _p]
        L.gal_bqueue_dequeue.argtypes = [_p, _p, _p, _u64, ctypes.c_int, _p]
        L.gal_bqueue_reclaim.argtypes = [_p, _p]
        L.gal_bqueue_stats.argtypes = [_p, _p, _p, _p]
        L.gal_init_global_host.argtypes = [_u64, _u64, ctypes.c_int]
        L.gal_global_host_handle.restype = _p
        L.gal_global_malloc_fused_bulk.argtypes = [_p, _u64, _p, _p]
        L.gal_global_free_fused_bulk.argtypes = [_p, _u64, _p]
        L.gal_global_malloc_combined_bulk.argtypes = [_p, _u64, ctypes.c_int, _p, _p]
        L.gal_global_free_combined_bulk.argtypes = [_p, _u64, ctypes.c_int, _p]
        L.gal_global_malloc_poison_bulk.argtypes = [_p, _u64, _p, _p]
        L.gal_global_check_poison_bulk.argtypes = [_p, _u64, _p, _p]
        L.gal_global_free_poison_bulk.argtypes = [_p, _u64, _p]
        L.gal_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise GallatinError(lib().gal_last_error().decode())
    return rc


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev_u64(n):
    import torch
    return torch.zeros(max(int(n), 1), dtype=torch.int64, device="cuda")


class Gallatin:
    """Gallatin<bytes_per_segment, smallest, biggest> (gallatin.cuh:292)."""

    def __init__(self, bytes_per_segment=16 << 20, smallest=16, biggest=4096):
        self.bytes_per_segment = bytes_per_segment
        self.smallest = smallest
        self.biggest = biggest
        self.handle = None

    # ---- host lifecycle (gallatin.cuh:324-421, 643-668, 1855) ----
    def generate_on_device(self, max_bytes, seed, print_info=True,
                           running_calloc=False, *, flags=0, slots0=128,
                           slots_min=4, limits=None, partitions=0):
        """limits: (request, slice, block, segment) retry limits; None or 0
        entries keep the reference's 1000, 5000, 500, 500.  partitions: XCD
        partitions (0 = one per XCD; deterministic mode uses 1)."""
        if running_calloc:
            raise GallatinError("calloc mode is not supported")
        cfg = Config()
        lib().gal_config_default(ctypes.byref(cfg), max_bytes, seed)
        cfg.bytes_per_segment = self.bytes_per_segment
        cfg.smallest = self.smallest
        cfg.biggest = self.biggest
        cfg.flags = flags
        cfg.slots0 = slots0
        cfg.slots_min = slots_min
        cfg.print_info = int(bool(print_info))
        if limits:
            (cfg.request_attempts, cfg.slice_attempts, cfg.block_attempts,
             cfg.segment_attempts) = [int(x or 0) for x in limits]
        cfg.partitions = partitions
        h = _p()
        _check(lib().gal_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h.value
        self.flags = flags
        self._info_static()
        return self

    def generate_on_device_host(self, max_bytes, seed, print_info=True, running_calloc=False,
                                **kw):
        """The pool in pinned, mapped host memory (gallatin.cuh:427-533)."""
        kw["flags"] = kw.get("flags", 0) | GAL_FLAG_POOL_HOST
        return self.generate_on_device(max_bytes, seed, print_info, running_calloc, **kw)

    def generate_on_device_managed(self, max_bytes, seed, print_info=True,
                                   running_calloc=False, **kw):
        """The pool in managed memory (gallatin.cuh:536-630)."""
        kw["flags"] = kw.get("flags", 0) | GAL_FLAG_POOL_MANAGED
        return self.generate_on_device(max_bytes, seed, print_info, running_calloc, **kw)

    def _info_static(self):
        i = self.info()
        self.num_segments = i["num_segments"]
        self.num_trees = i["num_trees"]
        self.blocks_per_segment = i["blocks_per_segment"]
        self.total_slots = i["total_slots"]
        self.partitions = i["partitions"]
        self.pool_base = i["pool_base"]

    def free_on_device(self):
        if self.handle:
            _check(lib().gal_destroy(self.handle))
            self.handle = None

    close = free_on_device

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free_on_device()

    def device_state(self):
        return lib().gal_device_state(self.handle)

    def print_info(self):
        _check(lib().gal_print_info(self.handle))

    STAT_NAMES = ("errors", "boot_failed", "segments_set_up", "segments_released",
                  "blocks_returned", "refills", "err_chunk_id", "err_ring_order",
                  "err_block_free", "err_ring_empty", "err_setup", "err_tag",
                  "err_multi_head", "err_invalid_free", "miss_no_block", "miss_retries",
                  "ring_runs", "ring_run_tickets")

    def stats(self):
        """Event counters (gal_get_stats) by name."""
        buf = (ctypes.c_uint64 * len(self.STAT_NAMES))()
        _check(lib().gal_get_stats(self.handle, buf, len(self.STAT_NAMES)))
        return {n: int(buf[i]) for i, n in enumerate(self.STAT_NAMES)}

    def info(self):
        i = Info()
        _check(lib().gal_get_info(self.handle, ctypes.byref(i)))
        arrays = ("tree_owned", "class_segments", "class_blocks_free", "class_blocks_used")
        d = {f: getattr(i, f) for f, _ in Info._fields_ if f not in arrays}
        for f in arrays:
            d[f] = list(getattr(i, f))[: i.num_trees]
        return d

    def drain(self):
        _check(lib().gal_drain(self.handle, _stream()))
        import torch
        torch.cuda.synchronize()

    def reboot(self):
        _check(lib().gal_reboot(self.handle, _stream()))

    def snapshot(self):
        """(seg_state u64[N], slots u32[total_slots], blocks u32[N*B, 2])."""
        n = self.num_segments
        st = np.zeros(n, np.uint64)
        sl = np.zeros(self.total_slots, np.uint32)
        bl = np.zeros(2 * n * self.blocks_per_segment, np.uint32)
        _check(lib().gal_snapshot(self.handle, st.ctypes.data, sl.ctypes.data,
                                  bl.ctypes.data))
        return st, sl, bl.reshape(-1, 2)

    def tree_layers(self, tree, partition=0):
        """Words of each layer (top first); tree -1 = segment tree, else the
        class tree of XCD partition `partition`."""
        nl = _u32()
        lw = (_u64 * 4)()
        n = _check(lib().gal_tree_words(self.handle, tree, partition, None, 0,
                                        ctypes.byref(nl), lw))
        out = np.zeros(n, np.uint64)
        _check(lib().gal_tree_words(self.handle, tree, partition, out.ctypes.data, n,
                                    ctypes.byref(nl), lw))
        layers, off = [], 0
        for layer in range(nl.value):
            layers.append(out[off:off + lw[layer]].copy())
            off += lw[layer]
        return layers

    # ---- workload kernels (tests/src/*.cu) ----
    def malloc_bulk(self, sizes, block_threads=256):
        """sizes: int64 cuda tensor -> int64 cuda tensor of device pointers."""
        import torch
        out = torch.zeros(sizes.numel(), dtype=torch.int64, device="cuda")
        _check(lib().gal_malloc_bulk(self.handle, sizes.data_ptr(), sizes.numel(),
                                     out.data_ptr(), block_threads, _stream()))
        return out

    def free_bulk(self, ptrs, block_threads=256):
        _check(lib().gal_free_bulk(self.handle, ptrs.data_ptr(), ptrs.numel(),
                                   block_threads, _stream()))

    def run_wave_stream(self, steps):
        """Replay wave steps (STEP_DTYPE) on one wave; byte offsets or FAIL."""
        import torch
        steps = np.ascontiguousarray(steps, dtype=STEP_DTYPE)
        raw = torch.from_numpy(steps.view(np.uint8).copy()).cuda()
        res = torch.zeros(len(steps) * 64, dtype=torch.int64, device="cuda")
        _check(lib().gal_run_wave_stream(self.handle, raw.data_ptr(), len(steps),
                                         res.data_ptr(), _stream()))
        torch.cuda.synchronize()
        return res.cpu().numpy().view(np.uint64)

    def run_plain(self, nthreads, size, counters=None, block_threads=256):
        ctr = counters if counters is not None else _dev_u64(CTR_COUNT)
        _check(lib().gal_run_plain(self.handle, nthreads, size, ctr.data_ptr(),
                                   block_threads, _stream()))
        return ctr

    def run_churn(self, seed, nthreads, rounds, counters=None, block_threads=256):
        ctr = counters if counters is not None else _dev_u64(CTR_COUNT)
        _check(lib().gal_run_churn(self.handle, seed, nthreads, rounds,
                                   ctr.data_ptr(), block_threads, _stream()))
        return ctr

    def run_churn_word(self, seed, nthreads, rounds, counters=None, batched=False,
                       workgroups=0, threads=0):
        """gal_run_churn_word: the churn stream with one word of each
        allocation written and checked (the reference's churn test pattern),
        per thread or through workgroup batches."""
        ctr = counters if counters is not None else _dev_u64(CTR_COUNT)
        _check(lib().gal_run_churn_word(self.handle, seed, nthreads, rounds, ctr.data_ptr(),
                                        1 if batched else 0, workgroups, threads, _stream()))
        return ctr

    def run_big(self, seed, step, n_small=4096, n_large=64, ptrs=None, counters=None):
        """One step of the block and multi-segment batch: n_small requests in
        [64 KiB, 64 MiB] and n_large in [256 MiB, 2 GiB] malloc and tag, then
        are checked and freed in a pseudo-random order.  Returns (ptrs,
        counters); ptrs keeps what the malloc launch returned."""
        import torch
        n = n_small + n_large
        if ptrs is None:
            ptrs = torch.zeros(n, dtype=torch.int64, device="cuda")
        if ptrs.numel() < n or ptrs.dtype != torch.int64 or not ptrs.is_cuda:
            raise ValueError("ptrs must be a device int64 tensor of n_small + n_large entries")
        ctr = counters if counters is not None else _dev_u64(CTR_COUNT)
        _check(lib().gal_run_big(self.handle, seed, step, n_small, n_large, ptrs.data_ptr(),
                                 ctr.data_ptr(), _stream()))
        return ptrs, ctr

    def run_churn_batched(self, seed, nthreads, rounds, counters=None, workgroups=0,
                          threads=0):
        ctr = counters if counters is not None else _dev_u64(CTR_COUNT)
        _check(lib().gal_run_churn_batched(self.handle, seed, nthreads, rounds,
                                           ctr.data_ptr(), workgroups, threads, _stream()))
        return ctr


def check_batch_stream(ops, offsets, nslots, persistent=True):
    """gal_check_batch_stream on host arrays: raises GallatinError naming
    the first op that breaks the stream rules (a slot filled again before
    its free, a free without an earlier malloc, and with persistent=True a
    slot filled twice in the stream)."""
    ops = np.ascontiguousarray(ops, dtype=BATCH_OP_DTYPE)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    _check(lib().gal_check_batch_stream(ops.ctypes.data_as(_p), offsets.ctypes.data_as(_p),
                                        len(offsets) - 1, int(nslots), int(bool(persistent))))


class BatchStream:
    """A batched request stream resident on the device: ops (BATCH_OP_DTYPE),
    batch offsets (host), and the pointer table the ops address."""

    def __init__(self, ops, offsets, nslots):
        import torch
        ops = np.ascontiguousarray(ops, dtype=BATCH_OP_DTYPE)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        self.nbatches = len(self.offsets) - 1
        self.nslots = int(nslots)
        self.n_ops = len(ops)
        self.n_mallocs = int((ops["kind"] == 0).sum())
        raw = torch.from_numpy(ops.view(np.uint8).copy())
        self.ops = raw.cuda() if len(ops) else torch.zeros(16, dtype=torch.uint8, device="cuda")
        self.ptrs = _dev_u64(self.nslots)

    def _offsets_ptr(self):
        return self.offsets.ctypes.data_as(ctypes.c_void_p)


BATCH_PLAIN, BATCH_WORKGROUP, BATCH_CACHE = 0, 1, 2


def _batch_mode(batched):
    """batched: False / 0 per-thread, True / 1 workgroup batches, 2 or
    "cache" mallocs through the workgroups' LDS slice caches."""
    if batched == "cache":
        return BATCH_CACHE
    mode = int(batched)
    if mode not in (BATCH_PLAIN, BATCH_WORKGROUP, BATCH_CACHE):
        raise ValueError("batched is 0, 1, 2 (or False, True, 'cache')")
    return mode


class BatchPlan:
    """gal_batch_plan: a stream's launches captured in one hipGraph."""

    def __init__(self, alloc, stream, counters, threads=0, batched=True, workgroups=0,
                 timed=False):
        """timed: the stream kernel also adds its wave time split to
        counters 7, 8, 10 and 11 (GAL_BATCH_TIMED; slows a replay)."""
        self.stream = stream  # keeps the buffers alive
        self.handle = _p()
        _check(lib().gal_batch_plan_create(
            alloc.handle, stream.ops.data_ptr(), stream._offsets_ptr(), stream.nbatches,
            stream.ptrs.data_ptr(), stream.nslots, counters.data_ptr(), threads,
            _batch_mode(batched) | (GAL_BATCH_TIMED if timed else 0), workgroups,
            ctypes.byref(self.handle)))

    def launch(self):
        _check(lib().gal_batch_plan_launch(self.handle, _stream()))

    def destroy(self):
        if self.handle:
            lib().gal_batch_plan_destroy(self.handle)
            self.handle = _p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def run_stream_static(alloc, stream, counters=None, threads=0, batched=True, workgroups=0,
                      timed=False):
    """gal_run_stream_static: the persistent form of the stream, run once on
    the current stream with the library's static workspace (no device
    allocation)."""
    ctr = counters if counters is not None else _dev_u64(CTR_COUNT)
    _check(lib().gal_run_stream_static(alloc.handle, stream.ops.data_ptr(), stream._offsets_ptr(),
                                       stream.nbatches, stream.ptrs.data_ptr(), stream.nslots,
                                       ctr.data_ptr(), threads,
                                       _batch_mode(batched) | (GAL_BATCH_TIMED if timed else 0),
                                       workgroups, _stream()))
    return ctr


def run_batches(alloc, stream, counters=None, threads=0, batched=True, workgroups=0):
    """Enqueue every batch of `stream` and the cleanup (gal_run_batches)."""
    ctr = counters if counters is not None else _dev_u64(CTR_COUNT)
    _check(lib().gal_run_batches(alloc.handle, stream.ops.data_ptr(), stream._offsets_ptr(),
                                 stream.nbatches, stream.ptrs.data_ptr(), stream.nslots,
                                 ctr.data_ptr(), threads, _batch_mode(batched), workgroups,
                                 _stream()))
    return ctr


# ---- allocator-backed data structures (on the global allocator) ----
class Vector:
    """fixed_vector<uint64_t, 64, 2^22> (fixed_vector.cuh:72-512)."""

    def __init__(self):
        self.handle = _p()
        _check(lib().gal_vector_create(ctypes.byref(self.handle)))

    def insert(self, items, bulk=False):
        import torch
        out = torch.empty(items.numel(), dtype=torch.int64, device="cuda")
        _check(lib().gal_vector_insert(self.handle, items.data_ptr(), items.numel(),
                                       out.data_ptr(), int(bool(bulk)), _stream()))
        return out

    def export(self):
        n = _check(lib().gal_vector_export(self.handle, None, 0))
        buf = np.zeros(max(n, 1), np.uint64)
        _check(lib().gal_vector_export(self.handle, buf.ctypes.data_as(ctypes.c_void_p), n))
        return buf[:n]

    def destroy(self):
        if self.handle:
            _check(lib().gal_vector_destroy(self.handle))
            self.handle = _p()


class Table:
    """extendible_hash_table of uint64 keys / values (extendible_ht.cuh:412);
    kind 0: 2^4..2^20 buckets, kind 1: 2^16..2^28."""

    def __init__(self, kind=0, seed=42):
        self.handle = _p()
        _check(lib().gal_table_create(kind, seed, ctypes.byref(self.handle)))

    def insert(self, keys, vals):
        import torch
        ok = torch.zeros(keys.numel(), dtype=torch.uint8, device="cuda")
        _check(lib().gal_table_insert(self.handle, keys.data_ptr(), vals.data_ptr(),
                                      keys.numel(), ok.data_ptr(), _stream()))
        return ok

    def query(self, keys):
        import torch
        vals = torch.zeros(keys.numel(), dtype=torch.int64, device="cuda")
        found = torch.zeros(keys.numel(), dtype=torch.uint8, device="cuda")
        _check(lib().gal_table_query(self.handle, keys.data_ptr(), keys.numel(),
                                     vals.data_ptr(), found.data_ptr(), _stream()))
        return vals, found

    def remove(self, keys):
        import torch
        ok = torch.zeros(keys.numel(), dtype=torch.uint8, device="cuda")
        _check(lib().gal_table_remove(self.handle, keys.data_ptr(), keys.numel(),
                                      ok.data_ptr(), _stream()))
        return ok

    def stats(self):
        v = [ctypes.c_uint64() for _ in range(4)]
        _check(lib().gal_table_stats(self.handle, *[ctypes.byref(x) for x in v]))
        return dict(zip(("keys", "buckets", "level", "failed"), (x.value for x in v)))

    def destroy(self):
        if self.handle:
            _check(lib().gal_table_destroy(self.handle))
            self.handle = _p()


class ChainTable:
    """chaining_table<Key, Val, size> (chaining_table.cuh:114-335) on the
    global allocator. Kinds (key/value width, pairs per block): 0 (64-bit,
    4), 1 (64-bit, 16), 2 (32-bit, 4), 3 (32-bit, 32), 4 (64-bit, 1).
    Keys and values are int64 tensors for 64-bit kinds, int32 for 32-bit."""

    KINDS = {0: (8, 4), 1: (8, 16), 2: (4, 4), 3: (4, 32), 4: (8, 1)}

    def __init__(self, kind=0, nslots=1 << 20, default_key=0, seed=123456, prealloc=False):
        if kind not in self.KINDS:
            raise GallatinError(f"unknown chaining table kind {kind}")
        self.kind = kind
        self.key_bytes, self.block_pairs = self.KINDS[kind]
        self.nblocks = nslots // self.block_pairs
        self.handle = _p()
        _check(lib().gal_chain_create(kind, nslots, default_key, seed, int(prealloc),
                                      ctypes.byref(self.handle)))

    def _dtype(self):
        import torch
        return torch.int64 if self.key_bytes == 8 else torch.int32

    def _check_arg(self, t):
        if t.dtype != self._dtype() or not t.is_cuda or not t.is_contiguous():
            raise GallatinError(f"expected a contiguous {self._dtype()} tensor on the GPU")

    def insert(self, keys, vals):
        import torch
        self._check_arg(keys)
        self._check_arg(vals)
        if keys.numel() != vals.numel():
            raise GallatinError("keys and values differ in length")
        ok = torch.zeros(keys.numel(), dtype=torch.uint8, device="cuda")
        _check(lib().gal_chain_insert(self.handle, keys.data_ptr(), vals.data_ptr(),
                                      keys.numel(), ok.data_ptr(), _stream()))
        return ok

    def query(self, keys):
        import torch
        self._check_arg(keys)
        vals = torch.zeros(keys.numel(), dtype=self._dtype(), device="cuda")
        found = torch.zeros(keys.numel(), dtype=torch.uint8, device="cuda")
        _check(lib().gal_chain_query(self.handle, keys.data_ptr(), keys.numel(),
                                     vals.data_ptr(), found.data_ptr(), _stream()))
        return vals, found

    def lengths(self):
        import torch
        out = torch.zeros(self.nblocks, dtype=torch.int32, device="cuda")
        _check(lib().gal_chain_lengths(self.handle, out.data_ptr(), _stream()))
        return out

    def stats(self):
        v = [ctypes.c_uint64() for _ in range(4)]
        _check(lib().gal_chain_stats(self.handle, *[ctypes.byref(x) for x in v]))
        return dict(zip(("max_chain", "blocks", "nblocks", "failed"), (x.value for x in v)))

    def print_chain_stats(self):
        _check(lib().gal_chain_print(self.handle))

    def destroy(self):
        if self.handle:
            _check(lib().gal_chain_destroy(self.handle))
            self.handle = _p()


class BlockQueue:
    """block_queue<uint64_t, items_per_block> (block_queue.cuh:240-463) on the
    global allocator; kind 0: 256 items per node, kind 1: 16."""

    ITEMS = {0: 256, 1: 16}

    def __init__(self, kind=0):
        if kind not in self.ITEMS:
            raise GallatinError(f"unknown block queue kind {kind}")
        self.items_per_block = self.ITEMS[kind]
        self.handle = _p()
        _check(lib().gal_bqueue_create(kind, ctypes.byref(self.handle)))

    def enqueue(self, items, serial=False):
        import torch
        ok = torch.zeros(items.numel(), dtype=torch.uint8, device="cuda")
        _check(lib().gal_bqueue_enqueue(self.handle, items.data_ptr(), items.numel(),
                                        ok.data_ptr(), int(serial), _stream()))
        return ok

    def dequeue(self, n, serial=False):
        import torch
        out = torch.zeros(n, dtype=torch.int64, device="cuda")
        ok = torch.zeros(n, dtype=torch.uint8, device="cuda")
        _check(lib().gal_bqueue_dequeue(self.handle, out.data_ptr(), ok.data_ptr(), n,
                                        int(serial), _stream()))
        return out, ok

    def reclaim(self):
        v = ctypes.c_uint64()
        _check(lib().gal_bqueue_reclaim(self.handle, ctypes.byref(v)))
        return v.value

    def stats(self):
        v = [ctypes.c_uint64() for _ in range(3)]
        _check(lib().gal_bqueue_stats(self.handle, *[ctypes.byref(x) for x in v]))
        return dict(zip(("nodes", "retired", "failed"), (x.value for x in v)))

    def destroy(self):
        if self.handle:
            _check(lib().gal_bqueue_destroy(self.handle))
            self.handle = _p()


class Queue:
    """queue<uint64_t, Gallatin> (queue.cuh:59-192) on an allocator; ring=True
    for circular_CAS_queue<uint64_t, 8, ...> (queue.cuh:196-253)."""

    def __init__(self, alloc, ring=False):
        self.handle = _p()
        _check(lib().gal_queue_create(alloc.handle, int(bool(ring)), ctypes.byref(self.handle)))

    def enqueue(self, items, serial=False):
        import torch
        ok = torch.zeros(items.numel(), dtype=torch.uint8, device="cuda")
        _check(lib().gal_queue_enqueue(self.handle, items.data_ptr(), items.numel(),
                                       ok.data_ptr(), int(bool(serial)), _stream()))
        return ok

    def dequeue(self, n, serial=False):
        import torch
        out = torch.zeros(n, dtype=torch.int64, device="cuda")
        ok = torch.zeros(n, dtype=torch.uint8, device="cuda")
        _check(lib().gal_queue_dequeue(self.handle, out.data_ptr(), ok.data_ptr(), n,
                                       int(bool(serial)), _stream()))
        return out, ok

    def destroy(self):
        if self.handle:
            _check(lib().gal_queue_destroy(self.handle))
            self.handle = _p()


# ---- global allocator (global_allocator.cuh:41-97) ----
def init_global_allocator(num_bytes, seed, print_info=True, running_calloc=False):
    if running_calloc:
        raise GallatinError("calloc mode is not supported")
    _check(lib().gal_init_global(num_bytes, seed, int(bool(print_info))))


def free_global_allocator():
    _check(lib().gal_free_global())


def print_global_stats():
    _check(lib().gal_print_global_stats())


def global_allocator():
    """Non-owning view of the global instance (for info / snapshots)."""
    h = lib().gal_global_handle()
    if not h:
        raise GallatinError("no global allocator")
    g = Gallatin()
    g.handle = h
    g.flags = 0
    g._info_static()
    return g


def global_malloc_bulk(sizes, block_threads=256):
    import torch
    out = torch.zeros(sizes.numel(), dtype=torch.int64, device="cuda")
    _check(lib().gal_global_malloc_bulk(sizes.data_ptr(), sizes.numel(),
                                        out.data_ptr(), block_threads, _stream()))
    return out


def global_free_bulk(ptrs, block_threads=256):
    _check(lib().gal_global_free_bulk(ptrs.data_ptr(), ptrs.numel(),
                                      block_threads, _stream()))


# ---- the host-pool instance, combined, fused, poison (global_allocator.cuh:100-357) ----

def init_global_allocator_host(num_bytes, seed, print_info=True, running_calloc=False):
    if running_calloc:
        raise GallatinError("calloc mode is not supported")
    _check(lib().gal_init_global_host(num_bytes, seed, int(bool(print_info))))


def free_global_allocator_host():
    _check(lib().gal_free_global_host())


def print_global_stats_host():
    _check(lib().gal_print_global_stats_host())


def global_host_allocator():
    """Non-owning view of the host-pool global instance."""
    h = lib().gal_global_host_handle()
    if not h:
        raise GallatinError("no host allocator")
    g = Gallatin()
    g.handle = h
    g.flags = GAL_FLAG_POOL_HOST
    g._info_static()
    return g


def init_global_allocator_combined(num_bytes, host_bytes, seed, print_info=True,
                                   running_calloc=False):
    init_global_allocator(num_bytes, seed, print_info, running_calloc)
    init_global_allocator_host(host_bytes, seed, print_info, running_calloc)


def free_global_allocator_combined():
    free_global_allocator()
    free_global_allocator_host()


def print_global_stats_combined():
    print("Device Allocator:", flush=True)
    print_global_stats()
    print("Host Allocator:", flush=True)
    print_global_stats_host()


def _ptrs_out(n):
    import torch
    return torch.zeros(n, dtype=torch.int64, device="cuda")


def global_malloc_fused_bulk(sizes):
    out = _ptrs_out(sizes.numel())
    _check(lib().gal_global_malloc_fused_bulk(sizes.data_ptr(), sizes.numel(), out.data_ptr(),
                                              _stream()))
    return out


def global_free_fused_bulk(ptrs):
    _check(lib().gal_global_free_fused_bulk(ptrs.data_ptr(), ptrs.numel(), _stream()))


def global_malloc_combined_bulk(sizes, on_host=False):
    out = _ptrs_out(sizes.numel())
    _check(lib().gal_global_malloc_combined_bulk(sizes.data_ptr(), sizes.numel(),
                                                 int(bool(on_host)), out.data_ptr(), _stream()))
    return out


def global_free_combined_bulk(ptrs, on_host=False):
    _check(lib().gal_global_free_combined_bulk(ptrs.data_ptr(), ptrs.numel(),
                                               int(bool(on_host)), _stream()))


def global_malloc_poison_bulk(sizes):
    out = _ptrs_out(sizes.numel())
    _check(lib().gal_global_malloc_poison_bulk(sizes.data_ptr(), sizes.numel(), out.data_ptr(),
                                               _stream()))
    return out


def global_check_poison_bulk(ptrs):
    import torch
    ok = torch.zeros(ptrs.numel(), dtype=torch.uint8, device="cuda")
    _check(lib().gal_global_check_poison_bulk(ptrs.data_ptr(), ptrs.numel(), ok.data_ptr(),
                                              _stream()))
    return ok


def global_free_poison_bulk(ptrs):
    _check(lib().gal_global_free_poison_bulk(ptrs.data_ptr(), ptrs.numel(), _stream()))


def bench_atomic(words, span, n_ops, lanes_per_address=1, stride=1):
    _check(lib().gal_bench_atomic(words.data_ptr(), span, n_ops,
                                  lanes_per_address, stride, _stream()))


def bench_copy(src, dst, nbytes):
    _check(lib().gal_bench_copy(src.data_ptr(), dst.data_ptr(), nbytes, _stream()))


def bench_tree(alloc, tree, searches, waves):
    sink = _dev_u64(1)
    _check(lib().gal_bench_tree(alloc.handle, tree, searches, waves, sink.data_ptr(),
                                _stream()))


def bench_scatter(src, dst, nbytes, piece_bytes):
    _check(lib().gal_bench_scatter(src.data_ptr(), dst.data_ptr(), nbytes, piece_bytes,
                                   _stream()))

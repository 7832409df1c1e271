"""TEST INFRASTRUCTURE ONLY - ctypes binding of the oracle (liboracle.so).

Only the tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` use this module, and only as the checker.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
FAIL = (1 << 64) - 1

_u64 = ctypes.c_uint64
_p = ctypes.c_void_p


class WaveStep(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("pad", ctypes.c_uint32),
                ("mask", ctypes.c_uint64), ("arg", ctypes.c_uint64 * 64)]


class Geometry(ctypes.Structure):
    _fields_ = [("num_segments", _u64), ("bytes_per_segment", _u64),
                ("smallest", _u64), ("biggest", _u64),
                ("blocks_per_segment", _u64), ("num_trees", ctypes.c_uint32),
                ("smallest_bits", ctypes.c_uint32),
                ("total_slots", ctypes.c_uint32), ("pad", ctypes.c_uint32),
                ("errors", _u64)]


STEP_DTYPE = np.dtype([("kind", "<u4"), ("pad", "<u4"), ("mask", "<u8"),
                       ("arg", "<u8", (64,))])
assert STEP_DTYPE.itemsize == ctypes.sizeof(WaveStep)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build as _b
            _b.build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_clog2.restype = ctypes.c_int
        L.orc_clog2.argtypes = [_u64]
        L.orc_cfcll.restype = ctypes.c_int
        L.orc_cfcll.argtypes = [_u64]
        L.orc_murmur64a.restype = _u64
        L.orc_murmur64a.argtypes = [_p, ctypes.c_int, _u64]
        L.orc_chain_of.restype = None
        L.orc_chain_of.argtypes = [_p, ctypes.c_int, _u64, _u64, _u64, _p]
        L.orc_big_stream.restype = None
        L.orc_big_stream.argtypes = [_u64, ctypes.c_uint32, _u64, _u64, _p, _p]
        L.orc_veb_create.restype = _p
        L.orc_veb_create.argtypes = [_u64, ctypes.c_int]
        L.orc_veb_destroy.argtypes = [_p]
        L.orc_veb_num_layers.restype = ctypes.c_int
        L.orc_veb_num_layers.argtypes = [_p]
        L.orc_veb_layer_words.restype = _u64
        L.orc_veb_layer_words.argtypes = [_p, ctypes.c_int]
        L.orc_veb_layer_bits.restype = ctypes.POINTER(_u64)
        L.orc_veb_layer_bits.argtypes = [_p, ctypes.c_int]
        for n in ("orc_veb_insert", "orc_veb_remove", "orc_veb_query"):
            getattr(L, n).restype = ctypes.c_int
            getattr(L, n).argtypes = [_p, _u64]
        for n in ("orc_veb_successor", "orc_veb_gather_multiple"):
            getattr(L, n).restype = _u64
            getattr(L, n).argtypes = [_p, _u64]
        for n in ("orc_veb_find_first", "orc_veb_malloc_first", "orc_veb_fill"):
            getattr(L, n).restype = _u64
            getattr(L, n).argtypes = [_p]
        L.orc_veb_return_multiple.argtypes = [_p, _u64, _u64]
        L.orc_create.restype = _p
        L.orc_create.argtypes = [_u64, _u64, _u64, _u64, ctypes.c_uint32,
                                 ctypes.c_uint32]
        L.orc_destroy.argtypes = [_p]
        L.orc_set_limits.argtypes = [_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int]
        L.orc_malloc_wave.argtypes = [_p, _u64, _p, _p]
        L.orc_free_wave.argtypes = [_p, _u64, _p]
        L.orc_malloc.restype = _u64
        L.orc_malloc.argtypes = [_p, _u64]
        L.orc_free.argtypes = [_p, _u64]
        L.orc_run_stream.restype = ctypes.c_int
        L.orc_run_stream.argtypes = [_p, _p, _u64, _p]
        L.orc_drain.argtypes = [_p]
        L.orc_get_geometry.argtypes = [_p, ctypes.POINTER(Geometry)]
        L.orc_segment_tree.restype = _p
        L.orc_segment_tree.argtypes = [_p]
        L.orc_sub_tree.restype = _p
        L.orc_sub_tree.argtypes = [_p, ctypes.c_int]
        L.orc_segment_state.argtypes = [_p, _p, _p]
        L.orc_slots.argtypes = [_p, _p]
        L.orc_blocks.argtypes = [_p, _p]
        L.orc_run_churn_sequential.restype = _u64
        L.orc_run_churn_sequential.argtypes = [_p, _u64, _u64, _u64,
                                               ctypes.c_uint32, _p, _p]
        _lib = L
    return _lib


def murmur64a(data: bytes, seed: int) -> int:
    buf = ctypes.create_string_buffer(data, len(data))
    return lib().orc_murmur64a(buf, len(data), seed)


def chain_of(keys, seed: int, nchains: int):
    """Chain index of each key of a numpy array of uint32 / uint64 keys."""
    import numpy as np
    k = np.ascontiguousarray(keys)
    out = np.zeros(len(k), dtype=np.uint64)
    lib().orc_chain_of(k.ctypes.data, k.dtype.itemsize, len(k), seed, nchains, out.ctypes.data)
    return out


def big_stream(seed: int, step: int, n_small: int, n_large: int):
    """(sizes, free order) of one step of the block and multi-segment
    request stream, from the header the GPU kernels use."""
    import numpy as np
    n = n_small + n_large
    sizes = np.zeros(n, dtype=np.uint64)
    order = np.zeros(n, dtype=np.uint64)
    lib().orc_big_stream(seed, step, n_small, n_large, sizes.ctypes.data, order.ctypes.data)
    return sizes, order


def clog2(x: int) -> int:
    return lib().orc_clog2(x)


def cfcll(x: int) -> int:
    return lib().orc_cfcll(x)


class Veb:
    """Standalone vEB tree of the oracle (for brute-force checks)."""

    def __init__(self, universe, filled=True, _ptr=None, _owned=True):
        self._owned = _owned
        self.ptr = _ptr if _ptr is not None else lib().orc_veb_create(universe, int(filled))
        if not self.ptr:
            raise ValueError("bad universe")

    def __del__(self):
        if getattr(self, "_owned", False) and getattr(self, "ptr", None):
            lib().orc_veb_destroy(self.ptr)
            self.ptr = None

    @property
    def num_layers(self):
        return lib().orc_veb_num_layers(self.ptr)

    def layer(self, i):
        n = lib().orc_veb_layer_words(self.ptr, i)
        p = lib().orc_veb_layer_bits(self.ptr, i)
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def set_layer(self, i, words):
        n = lib().orc_veb_layer_words(self.ptr, i)
        p = lib().orc_veb_layer_bits(self.ptr, i)
        arr = np.ctypeslib.as_array(p, shape=(n,))
        arr[:] = np.asarray(words, dtype=np.uint64)

    def layers(self):
        return [self.layer(i) for i in range(self.num_layers)]

    def insert(self, x): return lib().orc_veb_insert(self.ptr, x)
    def remove(self, x): return lib().orc_veb_remove(self.ptr, x)
    def query(self, x): return lib().orc_veb_query(self.ptr, x)
    def successor(self, x): return lib().orc_veb_successor(self.ptr, x)
    def find_first(self): return lib().orc_veb_find_first(self.ptr)
    def malloc_first(self): return lib().orc_veb_malloc_first(self.ptr)
    def gather_multiple(self, n): return lib().orc_veb_gather_multiple(self.ptr, n)
    def return_multiple(self, s, n): lib().orc_veb_return_multiple(self.ptr, s, n)
    def fill(self): return lib().orc_veb_fill(self.ptr)


class Allocator:
    """Sequential replay of the allocator policy (deterministic mode)."""

    def __init__(self, max_bytes, bytes_per_segment=16 << 20, smallest=16,
                 biggest=4096, slots0=128, slots_min=4, limits=None):
        self.ptr = lib().orc_create(max_bytes, bytes_per_segment, smallest,
                                    biggest, slots0, slots_min)
        if not self.ptr:
            raise ValueError("oracle: invalid geometry or boot failed")
        if limits:
            lib().orc_set_limits(self.ptr, *limits)
        g = Geometry()
        lib().orc_get_geometry(self.ptr, ctypes.byref(g))
        self.num_segments = g.num_segments
        self.bytes_per_segment = g.bytes_per_segment
        self.smallest = g.smallest
        self.biggest = g.biggest
        self.blocks_per_segment = g.blocks_per_segment
        self.num_trees = g.num_trees
        self.total_slots = g.total_slots

    def close(self):
        if getattr(self, "ptr", None):
            lib().orc_destroy(self.ptr)
            self.ptr = None

    __del__ = close

    @property
    def errors(self):
        g = Geometry()
        lib().orc_get_geometry(self.ptr, ctypes.byref(g))
        return g.errors

    def malloc(self, size):
        r = lib().orc_malloc(self.ptr, size)
        return None if r == FAIL else r

    def free(self, off):
        lib().orc_free(self.ptr, FAIL if off is None else off)

    def malloc_wave(self, mask, sizes):
        s = np.zeros(64, np.uint64)
        s[:len(sizes)] = sizes
        out = np.zeros(64, np.uint64)
        lib().orc_malloc_wave(self.ptr, mask, s.ctypes.data, out.ctypes.data)
        return out

    def free_wave(self, mask, offsets):
        o = np.full(64, FAIL, np.uint64)
        o[:len(offsets)] = offsets
        lib().orc_free_wave(self.ptr, mask, o.ctypes.data)

    def run_stream(self, steps: np.ndarray) -> np.ndarray:
        steps = np.ascontiguousarray(steps, dtype=STEP_DTYPE)
        res = np.zeros(len(steps) * 64, np.uint64)
        rc = lib().orc_run_stream(self.ptr, steps.ctypes.data, len(steps),
                                  res.ctypes.data)
        if rc != 0:
            raise ValueError("malformed stream")
        return res

    def drain(self):
        lib().orc_drain(self.ptr)

    def segment_tree(self):
        return Veb(0, _ptr=lib().orc_segment_tree(self.ptr), _owned=False)

    def sub_tree(self, t):
        return Veb(0, _ptr=lib().orc_sub_tree(self.ptr, t), _owned=False)

    def segment_state(self):
        ids = np.zeros(self.num_segments, np.uint16)
        act = np.zeros(self.num_segments, np.int32)
        lib().orc_segment_state(self.ptr, ids.ctypes.data, act.ctypes.data)
        return ids, act

    def slots(self):
        s = np.zeros(self.total_slots, np.uint32)
        lib().orc_slots(self.ptr, s.ctypes.data)
        return s

    def blocks(self):
        b = np.zeros(2 * self.num_segments * self.blocks_per_segment, np.uint32)
        lib().orc_blocks(self.ptr, b.ctypes.data)
        return b.reshape(-1, 2)

    def run_churn_sequential(self, seed, tid0, nthreads, rounds):
        miss = ctypes.c_uint64(0)
        bad = ctypes.c_uint64(0)
        pairs = lib().orc_run_churn_sequential(self.ptr, seed, tid0, nthreads,
                                               rounds, ctypes.byref(miss),
                                               ctypes.byref(bad))
        return pairs, miss.value, bad.value


def fresh_tree_layers(universe, filled):
    """Layers of a freshly initialised tree: all bits below each layer's
    universe set (segment tree) or all zero (size-class trees)."""
    t = Veb(universe, filled=filled)
    return t.layers()

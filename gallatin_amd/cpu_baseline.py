"""Host baselines: the device workloads' request streams through glibc
malloc/free on host threads (gallatin_amd/csrc/cpu_baseline.cpp, built as
libgallatin_cpu.so).  Each function returns the C result struct as a dict.
These measure the CPU side of the comparison; nothing on the device path
calls them."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgallatin_cpu.so")

_u64, _u32, _p, _f64 = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_double


class Result(ctypes.Structure):
    """gal_cpu_result"""
    _fields_ = [("pairs", _u64), ("misses", _u64), ("bad", _u64), ("bytes", _u64),
                ("seconds", _f64), ("threads", _u32), ("pad", _u32)]


class ServingResult(ctypes.Structure):
    """gal_cpu_serving_result"""
    _fields_ = [("appended_words", _u64), ("ended", _u64), ("table_ops", _u64),
                ("table_keys", _u64), ("bad", _u64), ("admitted", _u64), ("refused", _u64),
                ("held_bytes", _u64), ("live", _u32), ("threads", _u32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m gallatin_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.gal_cpu_churn_glibc.argtypes = [_u64, _u64, _u64, _u32, _u32, ctypes.POINTER(Result)]
        L.gal_cpu_big_glibc.argtypes = [_u64, _u32, _u32, _u64, _u64, _u32,
                                        ctypes.POINTER(Result)]
        L.gal_cpu_plain_glibc.argtypes = [_u64, _u32, _u32, ctypes.POINTER(Result)]
        L.gal_cpu_batches_glibc.argtypes = [_p, _p, _u32, _u64, _u32, ctypes.POINTER(Result)]
        L.gal_cpu_serving_glibc.argtypes = [_u64, _u32, _u32, _u32, _f64, _f64, _u32, _p, _u64,
                                            _f64, _u32, _p, ctypes.POINTER(ServingResult)]
        _lib = L
    return _lib


def _dict(r):
    return {f: getattr(r, f) for f, _ in r._fields_ if f != "pad"}


def _run(fn, *args):
    r = Result()
    if fn(*args, ctypes.byref(r)) != 0:
        raise RuntimeError(f"{fn.__name__} failed")
    return _dict(r)


def churn(seed, tid0, n, rounds, threads):
    return _run(lib().gal_cpu_churn_glibc, seed, tid0, n, rounds, threads)


def big(seed, step0, steps, n_small, n_large, threads):
    return _run(lib().gal_cpu_big_glibc, seed, step0, steps, n_small, n_large, threads)


def plain(n, launches, threads):
    return _run(lib().gal_cpu_plain_glibc, n, launches, threads)


def batches(ops, offsets, nslots, threads):
    ops = np.ascontiguousarray(ops)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    return _run(lib().gal_cpu_batches_glibc, ops.ctypes.data, offsets.ctypes.data,
                len(offsets) - 1, nslots, threads)


def serving(seed, warm, steps, max_sessions, arrivals_mean, p_end, ht_ops, tokens, pool_bytes,
            reserve_fraction, threads):
    """Returns (result dict, per-step milliseconds of the timed steps)."""
    tok = np.ascontiguousarray(tokens, dtype=np.uint32)
    ms = np.zeros(max(steps, 1))
    r = ServingResult()
    if lib().gal_cpu_serving_glibc(seed, warm, steps, max_sessions, arrivals_mean, p_end, ht_ops,
                                   tok.ctypes.data, pool_bytes, reserve_fraction, threads,
                                   ms.ctypes.data, ctypes.byref(r)) != 0:
        raise RuntimeError("gal_cpu_serving_glibc failed")
    return _dict(r), ms[:steps]

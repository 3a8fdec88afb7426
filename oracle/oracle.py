"""ctypes view of the CPU oracle (liboracle.so) and of oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's cpu_baseline leg, as the checker.  The product path never imports
this module.

dtype codes follow oracle.h: 0 float32, 1 float64, 2 uint32, 3 int32.
k values are 1-based ranks of the k-th largest element.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libggks_ref.so")

DTYPES = {np.dtype(np.float32): 0, np.dtype(np.float64): 1,
          np.dtype(np.uint32): 2, np.dtype(np.int32): 3}

_lib = None
_ref = None


def build(force: bool = False) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, i32, u32p = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p
        for name, args in {
            "ggko_sort": [i32, vp, i64, vp],
            "ggko_sort_choose": [i32, vp, i64, u32p, i64, vp],
            "ggko_topk": [i32, vp, i64, i64, vp],
            "ggko_radix_select": [i32, vp, i64, ctypes.c_uint32, vp, vp],
            "ggko_bucket_multiselect": [i32, vp, i64, u32p, i64, vp, ctypes.c_uint64, vp],
            "ggko_randomized_topk": [i32, vp, i64, i64, vp, ctypes.c_uint64],
            "ggko_nth_multiselect": [i32, vp, i64, u32p, i64, vp],
            "ggko_nth_topk": [i32, vp, i64, i64, vp],
            "ggko_parallel_sort_choose": [i32, vp, i64, u32p, i64, vp, i32],
            "ggko_keys": [i32, vp, vp, i64],
            "ggko_unkeys": [i32, vp, vp, i64],
        }.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int if name not in ("ggko_keys", "ggko_unkeys") else None
        L.ggko_key_f32.argtypes = [ctypes.c_uint32]
        L.ggko_key_f32.restype = ctypes.c_uint32
        L.ggko_unkey_f32.argtypes = [ctypes.c_uint32]
        L.ggko_unkey_f32.restype = ctypes.c_uint32
        L.ggko_key_f64.argtypes = [ctypes.c_uint64]
        L.ggko_key_f64.restype = ctypes.c_uint64
        L.ggko_unkey_f64.argtypes = [ctypes.c_uint64]
        L.ggko_unkey_f64.restype = ctypes.c_uint64
        _lib = L
    return _lib


class BucketStats(ctypes.Structure):
    _fields_ = [("num_buckets", ctypes.c_int64), ("total_count", ctypes.c_int64),
                ("reduced_length", ctypes.c_int64), ("unique_buckets", ctypes.c_int64),
                ("flat", ctypes.c_int32), ("monotone", ctypes.c_int32)]


def _prep(vec):
    v = np.ascontiguousarray(vec)
    if v.dtype not in DTYPES:
        raise TypeError(f"unsupported dtype {v.dtype}")
    return v, DTYPES[v.dtype]


def _ks(k):
    return np.ascontiguousarray(np.asarray(k, dtype=np.uint32))


def _check(rc, what):
    if rc != 0:
        raise ValueError(f"{what}: invalid argument (rc={rc})")


def keys(vec: np.ndarray) -> np.ndarray:
    """Converted (order-preserving unsigned) keys of vec."""
    v, dt = _prep(vec)
    out = np.empty(v.shape, np.uint64 if dt == 1 else np.uint32)
    lib().ggko_keys(dt, v.ctypes.data, out.ctypes.data, v.size)
    return out


def unkeys(keys_: np.ndarray, dtype) -> np.ndarray:
    dtype = np.dtype(dtype)
    k = np.ascontiguousarray(keys_, np.uint64 if dtype == np.float64 else np.uint32)
    out = np.empty(k.shape, dtype)
    lib().ggko_unkeys(DTYPES[dtype], k.ctypes.data, out.ctypes.data, k.size)
    return out


def sort(vec):
    v, dt = _prep(vec)
    out = np.empty_like(v)
    _check(lib().ggko_sort(dt, v.ctypes.data, v.size, out.ctypes.data), "sort")
    return out


def sort_choose(vec, k):
    v, dt = _prep(vec)
    k = _ks(k)
    out = np.empty(k.size, v.dtype)
    _check(lib().ggko_sort_choose(dt, v.ctypes.data, v.size, k.ctypes.data, k.size,
                                  out.ctypes.data), "sort_choose")
    return out


def topk(vec, K):
    v, dt = _prep(vec)
    out = np.empty(int(K), v.dtype)
    _check(lib().ggko_topk(dt, v.ctypes.data, v.size, int(K), out.ctypes.data), "topk")
    return out


def radix_select(vec, k):
    v, dt = _prep(vec)
    out = np.empty(1, v.dtype)
    digits = np.zeros(16, np.uint8)
    _check(lib().ggko_radix_select(dt, v.ctypes.data, v.size, int(k), out.ctypes.data,
                                   digits.ctypes.data), "radix_select")
    nd = 16 if dt == 1 else 8
    return out[0], digits[:nd]


def bucket_multiselect(vec, k, seed=0):
    v, dt = _prep(vec)
    k = _ks(k)
    out = np.empty(k.size, v.dtype)
    st = BucketStats()
    _check(lib().ggko_bucket_multiselect(dt, v.ctypes.data, v.size, k.ctypes.data, k.size,
                                         out.ctypes.data, seed, ctypes.byref(st)),
           "bucket_multiselect")
    return out, st


def randomized_topk(vec, K, seed=0):
    v, dt = _prep(vec)
    out = np.empty(int(K), v.dtype)
    _check(lib().ggko_randomized_topk(dt, v.ctypes.data, v.size, int(K), out.ctypes.data,
                                      seed), "randomized_topk")
    return out


def nth_multiselect(vec, k):
    v, dt = _prep(vec)
    k = _ks(k)
    out = np.empty(k.size, v.dtype)
    _check(lib().ggko_nth_multiselect(dt, v.ctypes.data, v.size, k.ctypes.data, k.size,
                                      out.ctypes.data), "nth_multiselect")
    return out


def nth_topk(vec, K):
    v, dt = _prep(vec)
    out = np.empty(int(K), v.dtype)
    _check(lib().ggko_nth_topk(dt, v.ctypes.data, v.size, int(K), out.ctypes.data),
           "nth_topk")
    return out


def parallel_sort_choose(vec, k, threads=0):
    v, dt = _prep(vec)
    k = _ks(k)
    out = np.empty(k.size, v.dtype)
    _check(lib().ggko_parallel_sort_choose(dt, v.ctypes.data, v.size, k.ctypes.data,
                                           k.size, out.ctypes.data, int(threads)),
           "parallel_sort_choose")
    return out


# ---------------------------------------------------------------- oracle/_ref
def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_PATH)
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        R.ggkr_sort_choose.argtypes = [i32, vp, i64, vp, i64, vp, i32,
                                       ctypes.POINTER(ctypes.c_double)]
        R.ggkr_sort_choose.restype = ctypes.c_int
        R.ggkr_topk.argtypes = [i32, vp, i64, i64, vp, i32, ctypes.POINTER(ctypes.c_double)]
        R.ggkr_topk.restype = ctypes.c_int
        _ref = R
    return _ref


def ref_sort_choose(vec, k, threads=0):
    """The reference's sort-and-choose through its vendored Thrust (OMP backend).
    Returns (values, sort_ms)."""
    v, dt = _prep(vec)
    k = _ks(k)
    out = np.empty(k.size, v.dtype)
    ms = ctypes.c_double(0.0)
    _check(ref().ggkr_sort_choose(dt, v.ctypes.data, v.size, k.ctypes.data, k.size,
                                  out.ctypes.data, int(threads), ctypes.byref(ms)),
           "ref_sort_choose")
    return out, ms.value


def ref_topk(vec, K, threads=0):
    v, dt = _prep(vec)
    out = np.empty(int(K), v.dtype)
    ms = ctypes.c_double(0.0)
    _check(ref().ggkr_topk(dt, v.ctypes.data, v.size, int(K), out.ctypes.data,
                           int(threads), ctypes.byref(ms)), "ref_topk")
    return out, ms.value

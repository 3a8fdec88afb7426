"""Loader for libggks.so (the HIP library) and its C ABI signatures.

The library is built in-tree by ``build.py`` (hipcc, gfx950).  There is no
fallback: if the shared object is missing, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libggks.so")

# ggks.h status codes
OK, EINVAL, ENOMEM, EDEVICE, EUNSUPPORTED = 0, -1, -2, -3, -4
F32, F64, U32, I32 = 0, 1, 2, 3
KIND_BUCKET, KIND_RADIX, KIND_TOPK = 0, 1, 2

# every symbol include/ggks.h declares (checked by the CPU test suite)
EXPORTS = [
    "ggks_ctx_create", "ggks_ctx_destroy", "ggks_scratch_bytes", "ggks_last_error",
    "ggks_version", "ggks_bucket_multiselect", "ggks_bucket_multiselect_f32",
    "ggks_bucket_multiselect_f64", "ggks_bucket_multiselect_u32",
    "ggks_bucket_multiselect_i32", "ggks_radix_multiselect", "ggks_topk",
    "ggks_radix_multiselect_f32", "ggks_radix_multiselect_f64", "ggks_radix_multiselect_u32",
    "ggks_radix_multiselect_i32", "ggks_topk_f32", "ggks_topk_f64", "ggks_topk_u32",
    "ggks_topk_i32",
    "ggks_bucket_multiselect_async", "ggks_radix_multiselect_async", "ggks_topk_async",
    "ggks_ctx_reserve", "ggks_ctx_check", "ggks_ctx_stream", "ggks_dev_alloc",
    "ggks_bucket_multiselect_noalloc", "ggks_sort_choose", "ggks_ctx_debug_regions",
    "ggks_dev_free", "ggks_h2d", "ggks_d2h", "ggks_d2d", "ggks_sync",
    "ggks_device_count", "ggks_set_device", "ggks_dev_fill", "ggks_ctx_profile_arm",
    "ggks_ctx_profile_read", "ggks_ctx_profile_marks", "ggks_batch_create", "ggks_batch_scratch_bytes", "ggks_batch_run",
    "ggks_batch_sync", "ggks_batch_stream", "ggks_batch_info", "ggks_batch_destroy",
    "ggks_tuning_set", "ggks_tuning_get", "ggks_tuning_reset", "ggks_generate",
    "ggks_tuning_single_set", "ggks_tuning_single_get",
    "ggks_server_create", "ggks_server_submit", "ggks_server_poll", "ggks_server_drain",
    "ggks_server_run_trace", "ggks_server_stats_get", "ggks_server_destroy",
]
BATCH_GRAPH = 1
GEN_UNIFORM, GEN_NORMAL, GEN_LEVELS, GEN_CAUCHY, GEN_PLATEAU = 0, 1, 2, 3, 4
GEN_SORTED, GEN_KILLER, GEN_NOISY, GEN_HALF_NORMAL, GEN_HUGE = 5, 6, 7, 8, 9
GEN_ONES, GEN_ONES_TWOS, GEN_NORMAL100 = 10, 11, 12
GEN_FAMILIES = {"uniform": 0, "normal": 1, "levels": 2, "cauchy": 3, "plateau": 4, "sorted": 5,
                "killer": 6, "noisy": 7, "half_normal": 8, "huge": 9, "ones": 10,
                "ones_twos": 11, "normal100": 12}
NSTAGES = 8
PROFILE_KERNEL_EVENTS = 0x80000000  # GGKS_PROFILE_KERNEL_EVENTS
NOALLOC_SCRATCH_BYTES = 64 << 20  # GGKS_NOALLOC_SCRATCH_BYTES
STAGES = ["queries", "sample", "lut", "count", "reduce", "findk", "compact", "finish"]


class GGKSError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ggks error {code}: {msg}")
        self.code = code


class Opts(ctypes.Structure):
    """ggks_opts (include/ggks.h)."""
    _fields_ = [("blocks", ctypes.c_int32), ("threads", ctypes.c_int32),
                ("buckets_log2", ctypes.c_int32), ("destructive", ctypes.c_int32),
                ("seed", ctypes.c_uint64), ("radix_bits", ctypes.c_int32),
                ("level2", ctypes.c_int32), ("reserved", ctypes.c_int32 * 4)]


class Segment(ctypes.Structure):
    """ggks_segment (include/ggks.h)."""
    _fields_ = [("offset", ctypes.c_uint64), ("length", ctypes.c_uint64),
                ("k_offset", ctypes.c_uint64), ("out_offset", ctypes.c_uint64),
                ("num_k", ctypes.c_uint32), ("kind", ctypes.c_uint32)]


SEGMENT_DTYPE = np.dtype([("offset", "<u8"), ("length", "<u8"), ("k_offset", "<u8"),
                          ("out_offset", "<u8"), ("num_k", "<u4"), ("kind", "<u4")])
assert SEGMENT_DTYPE.itemsize == ctypes.sizeof(Segment)


class BatchStats(ctypes.Structure):
    """ggks_batch_stats (include/ggks.h)."""
    _fields_ = [("elements", ctypes.c_uint64), ("queries", ctypes.c_uint64),
                ("scratch_bytes", ctypes.c_uint64), ("engine_segments", ctypes.c_uint32),
                ("direct_segments", ctypes.c_uint32), ("chunks", ctypes.c_uint32),
                ("graph", ctypes.c_uint32)]


REQUEST_DTYPE = np.dtype([("offset", "<u8"), ("length", "<u8"), ("h_k", "<u8"), ("h_out", "<u8"),
                          ("tag", "<u8"), ("num_k", "<u4"), ("kind", "<i4"), ("dtype", "<i4"),
                          ("reserved", "<i4")])


class Request(ctypes.Structure):
    """ggks_request (include/ggks.h)."""
    _fields_ = [("offset", ctypes.c_uint64), ("length", ctypes.c_uint64),
                ("h_k", ctypes.c_void_p), ("h_out", ctypes.c_void_p), ("tag", ctypes.c_uint64),
                ("num_k", ctypes.c_uint32), ("kind", ctypes.c_int32), ("dtype", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


assert REQUEST_DTYPE.itemsize == ctypes.sizeof(Request)


class DebugRegions(ctypes.Structure):
    """ggks_debug_regions (include/ggks.h)."""
    _fields_ = [(f, ctypes.c_void_p) for f in ("segs", "states", "lut", "chunks", "rows", "totals",
                                              "gsums", "bits", "children", "cq", "red")] + [
        ("n", ctypes.c_uint64), ("nchunks", ctypes.c_uint32), ("bmax", ctypes.c_uint32),
        ("nq", ctypes.c_uint32), ("pbits", ctypes.c_uint32), ("dtype", ctypes.c_int32),
        ("direct", ctypes.c_int32)]


class ServerStats(ctypes.Structure):
    """ggks_server_stats (include/ggks.h)."""
    _fields_ = [(f, ctypes.c_uint64) for f in (
        "submitted", "completed", "rejected", "batches", "elements", "max_pending", "pending",
        "scratch_budget", "scratch_in_use", "peak_scratch_bytes", "failed")]


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is not built; run `python __graft_entry__.py` or "
            "`python ggks-multiselect_amd/build.py` first")
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, i32, sz = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_size_t
    optp = ctypes.POINTER(Opts)
    sig = {
        "ggks_ctx_create": ([ctypes.POINTER(vp), i32, vp], i32),
        "ggks_ctx_destroy": ([vp], i32),
        "ggks_scratch_bytes": ([i32, i32, i64, i64], sz),
        "ggks_last_error": ([], ctypes.c_char_p),
        "ggks_version": ([], ctypes.c_char_p),
        "ggks_bucket_multiselect": ([vp, i32, vp, i64, vp, i64, vp, optp], i32),
        "ggks_radix_multiselect": ([vp, i32, vp, i64, vp, i64, vp, optp], i32),
        "ggks_topk": ([vp, i32, vp, i64, i64, vp, optp], i32),
        "ggks_bucket_multiselect_async": ([vp, i32, vp, i64, vp, i64, vp, optp], i32),
        "ggks_radix_multiselect_async": ([vp, i32, vp, i64, vp, i64, vp, optp], i32),
        "ggks_topk_async": ([vp, i32, vp, i64, i64, vp, optp], i32),
        "ggks_ctx_reserve": ([vp, i32, i32, i64, i64], i32),
        "ggks_bucket_multiselect_noalloc": ([i32, vp, i64, vp, i64, vp, optp, vp], i32),
        "ggks_sort_choose": ([vp, i32, vp, i64, vp, i64, vp], i32),
        "ggks_ctx_debug_regions": ([vp, ctypes.POINTER(DebugRegions)], i32),
        "ggks_ctx_check": ([vp], i32),
        "ggks_ctx_stream": ([vp], vp),
        "ggks_dev_alloc": ([ctypes.POINTER(vp), sz], i32),
        "ggks_dev_free": ([vp], i32),
        "ggks_h2d": ([vp, vp, sz], i32),
        "ggks_d2h": ([vp, vp, sz], i32),
        "ggks_d2d": ([vp, vp, sz], i32),
        "ggks_sync": ([], i32),
        "ggks_device_count": ([ctypes.POINTER(i32)], i32),
        "ggks_set_device": ([i32], i32),
        "ggks_dev_fill": ([vp, i32, sz], i32),
        "ggks_ctx_profile_arm": ([vp, i32], i32),
        "ggks_ctx_profile_read": ([vp, vp, i32], i32),
        "ggks_ctx_profile_marks": ([vp, ctypes.c_uint32], i32),
        "ggks_batch_create": ([ctypes.POINTER(vp), i32, i32, vp, i64, vp, i64, optp, vp, sz, vp],
                              i32),
        "ggks_batch_scratch_bytes": ([i32, vp, i64, vp, i64, optp], sz),
        "ggks_batch_run": ([vp, vp, vp, i32], i32),
        "ggks_batch_sync": ([vp], i32),
        "ggks_batch_stream": ([vp], vp),
        "ggks_batch_info": ([vp, ctypes.POINTER(BatchStats)], i32),
        "ggks_batch_destroy": ([vp], i32),
        "ggks_tuning_set": ([i32, i32, i32, i32, i32], i32),
        "ggks_tuning_get": ([i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i32),
                             ctypes.POINTER(i32)], i32),
        "ggks_tuning_reset": ([], i32),
        "ggks_tuning_single_set": ([i32, i32, i32, i32], i32),
        "ggks_tuning_single_get": ([i32, i32, ctypes.POINTER(i32), ctypes.POINTER(i32)], i32),
        "ggks_generate": ([vp, i32, i32, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                           i64, i64], i32),
        "ggks_server_create": ([ctypes.POINTER(vp), i32, vp, vp, sz, i32, i32], i32),
        "ggks_server_submit": ([vp, ctypes.POINTER(Request)], i32),
        "ggks_server_poll": ([vp, vp, vp, i32], i32),
        "ggks_server_drain": ([vp], i32),
        "ggks_server_run_trace": ([vp, vp, vp, i64, vp], i32),
        "ggks_server_stats_get": ([vp, ctypes.POINTER(ServerStats)], i32),
        "ggks_server_destroy": ([vp], i32),
    }
    for t in ("f32", "f64", "u32", "i32"):
        sig[f"ggks_bucket_multiselect_{t}"] = ([vp, i64, vp, ctypes.c_int32, vp, optp], i32)
        sig[f"ggks_radix_multiselect_{t}"] = ([vp, i64, vp, ctypes.c_int32, vp, optp], i32)
        sig[f"ggks_topk_{t}"] = ([vp, i64, ctypes.c_int32, vp, optp], i32)
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != OK:
        msg = lib().ggks_last_error()
        raise GGKSError(rc, msg.decode() if msg else "")

"""Python face of libggks.so, mirroring the reference's wrapper interface.

Reference entry points and their counterparts here:

=============================================================  ==================================
reference                                                      here
=============================================================  ==================================
BucketMultiselect::bucketMultiselectWrapper                    :func:`bucketMultiselectWrapper`
  (bucketMultiselect.cu:837-871)
RadixSelect::RadixSelectWrapper (merrillSelect.cu:210-259)     :func:`RadixSelectWrapper`
randomizedTopkSelectWrapper (randomizedTopkSelect.cu:210-213)  :func:`randomizedTopkSelectWrapper`
timeBucketTopkselect's k = 1..K call                           :func:`topk`
  (multiselectTimingFunctions.cu:204-231)
=============================================================  ==================================

Semantics follow the reference: k is the 1-based rank of the k-th largest
element, outputs follow the caller's k order, top-K is descending.  Unlike the
reference, the input vector is never modified and errors raise
:class:`GGKSError` instead of printing and returning the success value
(bucketMultiselect.cu:30-32).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import GGKSError, Opts, check, lib

_DTYPE_CODE = {np.dtype(np.float32): _lib.F32, np.dtype(np.float64): _lib.F64,
               np.dtype(np.uint32): _lib.U32, np.dtype(np.int32): _lib.I32}


def dtype_code(dtype) -> int:
    d = np.dtype(dtype)
    if d not in _DTYPE_CODE:
        raise TypeError(f"unsupported dtype {d}; use float32, float64, uint32 or int32")
    return _DTYPE_CODE[d]


class DeviceArray:
    """A device buffer owned by this process (allocated through libggks)."""

    def __init__(self, n: int, dtype):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = self.n * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().ggks_dev_alloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if a.size:
            check(lib().ggks_h2d(d.ptr, a.ctypes.data, d.nbytes))
        return d

    def to_host(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        if self.n:
            check(lib().ggks_d2h(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def view(self, offset: int, n: int) -> "DeviceView":
        return DeviceView(self, offset, n)

    def free(self) -> None:
        if self.ptr:
            lib().ggks_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __len__(self):
        return self.n


class DeviceView:
    """A window [offset, offset + n) of a DeviceArray (no ownership)."""

    def __init__(self, base: DeviceArray, offset: int, n: int):
        if offset < 0 or n < 0 or offset + n > base.n:
            raise ValueError("view out of range")
        self.base, self.dtype, self.n = base, base.dtype, int(n)
        self.ptr = base.ptr + offset * base.dtype.itemsize
        self.nbytes = self.n * self.dtype.itemsize

    def __len__(self):
        return self.n


class Context:
    """A ggks_ctx: its own stream (None = null stream) and scratch arena."""

    def __init__(self, device: int = 0, stream: int | None = None):
        p = ctypes.c_void_p()
        check(lib().ggks_ctx_create(ctypes.byref(p), int(device), stream))
        self.handle = p.value

    def close(self):
        if self.handle:
            lib().ggks_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _opts(blocks=0, threads=0, buckets_log2=0, seed=0, radix_bits=0, level2=0):
    o = Opts()
    o.blocks, o.threads, o.buckets_log2 = int(blocks), int(threads), int(buckets_log2)
    o.seed, o.radix_bits, o.level2 = int(seed), int(radix_bits), int(level2)
    return o


def _k_array(kvals) -> np.ndarray:
    k = np.asarray(kvals)
    if k.size and (k.min() < 1 or k.max() > 0xFFFFFFFF):
        raise GGKSError(_lib.EINVAL, "k outside [1, 2^32)")
    return np.ascontiguousarray(k, dtype=np.uint32)


def multiselect(dvec, kvals, ctx: Context | None = None, kind: str = "bucket",
                **opt) -> np.ndarray:
    """Values of the k-th largest elements of device vector `dvec`, in k order."""
    k = _k_array(kvals)
    out = np.empty(k.size, dvec.dtype)
    fn = lib().ggks_bucket_multiselect if kind == "bucket" else lib().ggks_radix_multiselect
    o = _opts(**opt)
    check(fn(ctx.handle if ctx else None, dtype_code(dvec.dtype), dvec.ptr, len(dvec),
             k.ctypes.data, k.size, out.ctypes.data, ctypes.byref(o)))
    return out


def topk(dvec, K: int, ctx: Context | None = None, **opt) -> np.ndarray:
    """The K largest values of `dvec`, descending."""
    out = np.empty(int(K), dvec.dtype)
    o = _opts(**opt)
    check(lib().ggks_topk(ctx.handle if ctx else None, dtype_code(dvec.dtype), dvec.ptr,
                          len(dvec), int(K), out.ctypes.data, ctypes.byref(o)))
    return out


def multiselect_async(dvec, d_k: DeviceArray, d_out: DeviceArray, ctx: Context,
                      kind: str = "bucket", **opt) -> None:
    """Stream-ordered select: device k list in, device outputs out."""
    fn = (lib().ggks_bucket_multiselect_async if kind == "bucket"
          else lib().ggks_radix_multiselect_async)
    o = _opts(**opt)
    check(fn(ctx.handle, dtype_code(dvec.dtype), dvec.ptr, len(dvec), d_k.ptr, len(d_k),
             d_out.ptr, ctypes.byref(o)))


def scratch_bytes(kind: str, dtype, n: int, num_k: int) -> int:
    code = {"bucket": _lib.KIND_BUCKET, "radix": _lib.KIND_RADIX, "topk": _lib.KIND_TOPK}[kind]
    return int(lib().ggks_scratch_bytes(code, dtype_code(dtype), int(n), int(num_k)))


def sync() -> None:
    check(lib().ggks_sync())


# ----------------------------------------------- reference-named wrappers
def bucketMultiselectWrapper(d_vector, length: int, kVals_ori, numKs: int, outputs,
                             blocks: int = 0, threads: int = 0):
    """bucketMultiselect.cu:837-871: fills host array `outputs[numKs]` with the
    kVals_ori[i]-th largest values of the first `length` elements of
    `d_vector`.  Returns 1 as the reference does."""
    if length < 1 or length > len(d_vector):
        raise GGKSError(_lib.EINVAL, "length outside [1, len(d_vector)]")
    vec = d_vector if length == len(d_vector) else DeviceView(
        d_vector if isinstance(d_vector, DeviceArray) else d_vector.base, 0, length)
    k = _k_array(np.asarray(kVals_ori)[:numKs])
    res = multiselect(vec, k, blocks=blocks, threads=threads)
    outputs[:numKs] = res
    return 1


def RadixSelectWrapper(d_vec, size: int, k: int):
    """merrillSelect.cu:210-259: the k-th largest of the first `size`
    elements (input left intact)."""
    vec = d_vec if size == len(d_vec) else DeviceView(
        d_vec if isinstance(d_vec, DeviceArray) else d_vec.base, 0, size)
    return multiselect(vec, [k], kind="radix")[0]


def randomizedTopkSelectWrapper(lst, size: int, k: int) -> np.ndarray:
    """randomizedTopkSelect.cu:210-213: the k largest values, descending, as a
    new host array (the reference returns a malloc'd array)."""
    vec = lst if size == len(lst) else DeviceView(
        lst if isinstance(lst, DeviceArray) else lst.base, 0, size)
    return topk(vec, k)


# ------------------------------------------------------------ batched select
_KIND_CODE = {"bucket": _lib.KIND_BUCKET, "radix": _lib.KIND_RADIX, "topk": _lib.KIND_TOPK}


def segment_table(lengths, kinds, num_k, offsets=None, k_offsets=None,
                  out_offsets=None) -> np.ndarray:
    """A ggks_segment table (numpy structured array).  Offsets default to
    back-to-back segments, k lists and outputs."""
    lengths = np.asarray(lengths, np.uint64)
    num_k = np.asarray(num_k, np.uint32)
    kinds = np.asarray([_KIND_CODE[k] if isinstance(k, str) else int(k) for k in kinds], np.uint32)
    t = np.zeros(lengths.size, _lib.SEGMENT_DTYPE)
    t["length"] = lengths
    t["num_k"] = num_k
    t["kind"] = kinds
    t["offset"] = offsets if offsets is not None else np.concatenate(
        [[0], np.cumsum(lengths)[:-1]]).astype(np.uint64)
    nk_lists = np.where(kinds == _lib.KIND_TOPK, 0, num_k).astype(np.uint64)
    t["k_offset"] = k_offsets if k_offsets is not None else np.concatenate(
        [[0], np.cumsum(nk_lists)[:-1]]).astype(np.uint64)
    t["out_offset"] = out_offsets if out_offsets is not None else np.concatenate(
        [[0], np.cumsum(num_k.astype(np.uint64))[:-1]]).astype(np.uint64)
    return t


class Batch:
    """A planned ggks_batch: many segments of one device vector selected by
    one launch sequence (optionally a replayed hipGraph).  `table` is a
    segment_table(); `k` the concatenated k lists of its multiselect
    segments."""

    def __init__(self, dtype, table: np.ndarray, k=None, device: int = 0, scratch=None,
                 stream: int | None = None, **opt):
        self.dtype = np.dtype(dtype)
        self.table = np.ascontiguousarray(table, dtype=_lib.SEGMENT_DTYPE)
        self.k = np.ascontiguousarray(np.zeros(0, np.uint32) if k is None else k, np.uint32)
        o = _opts(**opt)
        p = ctypes.c_void_p()
        sp = scratch.ptr if scratch is not None else None
        sb = scratch.nbytes if scratch is not None else 0
        check(lib().ggks_batch_create(ctypes.byref(p), int(device), dtype_code(self.dtype),
                                      self.table.ctypes.data, self.table.size,
                                      self.k.ctypes.data, self.k.size, ctypes.byref(o), sp, sb,
                                      stream))
        self.handle = p.value
        self.num_outputs = int((self.table["out_offset"] + self.table["num_k"]).max()) \
            if self.table.size else 0

    @staticmethod
    def scratch_bytes(dtype, table: np.ndarray, k=None, **opt) -> int:
        table = np.ascontiguousarray(table, dtype=_lib.SEGMENT_DTYPE)
        k = np.ascontiguousarray(np.zeros(0, np.uint32) if k is None else k, np.uint32)
        o = _opts(**opt)
        return int(lib().ggks_batch_scratch_bytes(dtype_code(dtype), table.ctypes.data,
                                                  table.size, k.ctypes.data, k.size,
                                                  ctypes.byref(o)))

    def run(self, dvec, dout, graph: bool = False) -> None:
        check(lib().ggks_batch_run(self.handle, dvec.ptr, dout.ptr,
                                   _lib.BATCH_GRAPH if graph else 0))

    def sync(self) -> None:
        check(lib().ggks_batch_sync(self.handle))

    def stats(self) -> dict:
        st = _lib.BatchStats()
        check(lib().ggks_batch_info(self.handle, ctypes.byref(st)))
        return {f: int(getattr(st, f)) for f, _ in st._fields_}

    def __call__(self, dvec, graph: bool = False) -> np.ndarray:
        out = DeviceArray(max(self.num_outputs, 1), self.dtype)
        self.run(dvec, out, graph)
        self.sync()
        return out.to_host()[:self.num_outputs]

    def close(self):
        if self.handle:
            lib().ggks_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tuning_set(dtype, log2_len: int, buckets_log2: int, chunk_log2: int, direct: bool) -> None:
    check(lib().ggks_tuning_set(dtype_code(dtype), int(log2_len), int(buckets_log2),
                                int(chunk_log2), int(bool(direct))))


def tuning_get(dtype, log2_len: int) -> tuple:
    b, c, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    check(lib().ggks_tuning_get(dtype_code(dtype), int(log2_len), ctypes.byref(b),
                                ctypes.byref(c), ctypes.byref(d)))
    return b.value, c.value, bool(d.value)


def tuning_reset() -> None:
    check(lib().ggks_tuning_reset())


def tuning_single_set(dtype, log2_len: int, buckets_log2: int, chunks: int) -> None:
    check(lib().ggks_tuning_single_set(dtype_code(dtype), int(log2_len), int(buckets_log2),
                                       int(chunks)))


def tuning_single_get(dtype, log2_len: int) -> tuple:
    b, c = ctypes.c_int(), ctypes.c_int()
    check(lib().ggks_tuning_single_get(dtype_code(dtype), int(log2_len), ctypes.byref(b),
                                       ctypes.byref(c)))
    return b.value, c.value


def generate(dtype, family: str, seed: int, n: int, start: int = 0, lo: int = 0, hi: int = 0,
             out: DeviceArray | None = None, offset: int = 0):
    """Fill a device array (new, or `out` from element `offset`) with n seeded
    values of a family of ggks.h's ggks_generate: "uniform" (floats in [0,1),
    integers in [lo, hi]), "normal", "levels" (floats holding the integers
    lo..hi), "cauchy", "plateau", "sorted" (positions start.. of a
    non-decreasing sequence of length hi), "killer", "noisy", "half_normal",
    "huge", "ones", "ones_twos" or "normal100"."""
    d = out if out is not None else DeviceArray(n, dtype)
    fam = _lib.GEN_FAMILIES[family]
    if offset < 0 or offset + n > d.n:
        raise ValueError("generate range outside the array")
    check(lib().ggks_generate(d.ptr + offset * d.dtype.itemsize, dtype_code(d.dtype), fam,
                              int(seed) & 0xFFFFFFFFFFFFFFFF, int(start), int(n), int(lo), int(hi)))
    return d


# ------------------------------------------------------------------ serving
class Server:
    """A ggks_server: continuous batching of select requests over vectors
    resident in HBM.  `resident` maps a dtype to the DeviceArray holding its
    vectors; requests name a vector by (dtype, offset, length)."""

    def __init__(self, resident: dict, scratch_budget: int, max_inflight: int = 4,
                 max_batch: int = 256, device: int = 0):
        ptrs = (ctypes.c_void_p * 4)()
        lens = (ctypes.c_uint64 * 4)()
        self._resident = dict(resident)
        for dt, arr in resident.items():
            c = dtype_code(dt)
            if np.dtype(dt) != arr.dtype:
                raise TypeError("resident buffer dtype mismatch")
            ptrs[c], lens[c] = arr.ptr, arr.n
        p = ctypes.c_void_p()
        check(lib().ggks_server_create(ctypes.byref(p), int(device), ptrs, lens,
                                       int(scratch_budget), int(max_inflight), int(max_batch)))
        self.handle = p.value
        # internal tag -> (caller's tag, k, out): the host buffers a pending
        # request points at stay alive until it is reported; internal tags
        # are unique, so callers may reuse or repeat their own tags
        self._live = {}
        self._next = 1

    def submit(self, dtype, offset: int, length: int, out: np.ndarray, k=None, K: int = 0,
               kind: str = "bucket", tag: int = 0) -> None:
        r = _lib.Request()
        itag = self._next
        r.offset, r.length, r.tag = int(offset), int(length), itag
        r.dtype = dtype_code(dtype)
        if kind == "topk":
            r.kind, r.num_k, kk = _lib.KIND_TOPK, int(K), None
        else:
            kk = _k_array(k)
            r.kind, r.num_k, r.h_k = _KIND_CODE[kind], kk.size, kk.ctypes.data
        if out.dtype != np.dtype(dtype) or out.size < r.num_k or not out.flags.c_contiguous:
            raise ValueError("output buffer must be a contiguous array of the dtype, num_k long")
        r.h_out = out.ctypes.data
        check(lib().ggks_server_submit(self.handle, ctypes.byref(r)))
        self._next += 1
        self._live[itag] = (int(tag), kk, out)

    def poll(self, max_done: int = 4096) -> list:
        tags = np.zeros(max_done, np.uint64)
        lat = np.zeros(max_done, np.float64)
        n = lib().ggks_server_poll(self.handle, tags.ctypes.data, lat.ctypes.data, max_done)
        if n < 0:
            check(n)
        done = []
        for t, l in zip(tags[:n].tolist(), lat[:n].tolist()):
            user_tag = self._live.pop(int(t))[0]
            done.append((user_tag, l))
        return done

    def drain(self) -> None:
        check(lib().ggks_server_drain(self.handle))

    def run_trace(self, reqs: np.ndarray, arrival_s: np.ndarray) -> np.ndarray:
        """Replay an arrival trace (REQUEST_DTYPE records whose h_k / h_out
        point at live host buffers); returns latency per request (s, -1 if
        refused)."""
        reqs = np.ascontiguousarray(reqs, _lib.REQUEST_DTYPE)
        arrival_s = np.ascontiguousarray(arrival_s, np.float64)
        lat = np.full(reqs.size, -1.0)
        check(lib().ggks_server_run_trace(self.handle, reqs.ctypes.data, arrival_s.ctypes.data,
                                          reqs.size, lat.ctypes.data))
        return lat

    def stats(self) -> dict:
        st = _lib.ServerStats()
        check(lib().ggks_server_stats_get(self.handle, ctypes.byref(st)))
        return {f: int(getattr(st, f)) for f, _ in st._fields_}

    def close(self):
        if self.handle:
            lib().ggks_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["DeviceArray", "DeviceView", "Context", "GGKSError", "multiselect", "topk",
           "multiselect_async", "scratch_bytes", "sync", "bucketMultiselectWrapper",
           "RadixSelectWrapper", "randomizedTopkSelectWrapper", "dtype_code", "Batch",
           "segment_table", "tuning_set", "tuning_get", "tuning_reset", "generate", "Server",
           "tuning_single_set", "tuning_single_get"]

"""Host-side readers and a numpy restatement of the engine's bucket function,
for the device-invariant tests: they read the tables a select left in its
context (ggks_ctx_debug_regions, include/ggks.h) and check them against
properties the algorithm fixes (every key lands in exactly one bucket, the
bucket function is monotone in the key, the chunk rows are exclusive prefix
sums of per-chunk histograms, each child's reduced slice is the multiset of
its bucket's keys, and each child query's rank is the requested order
statistic's rank inside the child).

The restatement follows BucketFn in ggks-multiselect_amd/csrc/ggks_kernels.hpp
(split + finish + heavy); the record layouts are those of
ggks-multiselect_amd/csrc/ggks_types.hpp, which test_debug_layout compiles
and compares field by field.
"""
import ctypes

import numpy as np

SEGDESC = np.dtype([("data_off", "<u8"), ("len", "<u8"), ("row_base", "<u8"), ("red_base", "<u8"),
                    ("lo", "<u8"), ("hi", "<u8"), ("bmax", "<u4"), ("pbits", "<u4"),
                    ("nchunks", "<u4"), ("sample_off", "<u4"), ("nsample", "<u4"),
                    ("lut_off", "<u4"), ("tot_off", "<u4"), ("bits_off", "<u4"), ("q_off", "<u4"),
                    ("q_cnt", "<u4"), ("mode", "<u4"), ("radix_bits", "<u4"), ("qsrc", "<u4"),
                    ("out_off", "<u4"), ("flags", "<u4")], align=True)
SEGSTATE = np.dtype([("kb", "<u8"), ("kmax", "<u8"), ("lmax", "<u8"), ("hv", "<u8"), ("s", "<u4"),
                     ("P", "<u4"), ("B", "<u4"), ("nact", "<u4"), ("poff", "<u4"), ("pwid", "<u4"),
                     ("rsh", "<u4"), ("twid", "<u4"), ("lsh", "<u4"), ("hvon", "<u4"), ("hb", "<u4"),
                     ("nfree", "<u4"), ("arrive", "<u4"), ("nred", "<u4")], align=True)
CHUNK = np.dtype([("seg", "<u4"), ("idx", "<u4"), ("start", "<u8"), ("len", "<u8")], align=True)
CHILD = np.dtype([("start", "<u8"), ("len", "<u8"), ("lo", "<u8"), ("hi", "<u8"), ("q_off", "<u4"),
                  ("q_cnt", "<u4"), ("src", "<u4"), ("need_range", "<u4"), ("bucket", "<u4"),
                  ("seg", "<u4")], align=True)
QUERY = np.dtype([("rank", "<u4"), ("out", "<u4")])
REDUCE_BINS = 64  # kReduceBins: bins per group of the bucket prefixes


def read(lib, ptr: int, dtype, count: int) -> np.ndarray:
    """count records of dtype at device address ptr."""
    out = np.empty(count, dtype)
    if count:
        rc = lib.ggks_d2h(out.ctypes.data, ctypes.c_void_p(ptr), out.nbytes)
        assert rc == 0, f"ggks_d2h rc={rc}"
    return out


def regions(pkg, ctx):
    L = pkg._lib
    g = L.DebugRegions()
    assert L.lib().ggks_ctx_debug_regions(ctx.handle, ctypes.byref(g)) == 0
    return g


def to_key(x: np.ndarray) -> np.ndarray:
    """The order-preserving unsigned key of each element (ggks_keys.hpp to_key)."""
    if x.dtype == np.float32 or x.dtype == np.float64:
        u = np.uint32 if x.dtype == np.float32 else np.uint64
        w = 32 if u is np.uint32 else 64
        b = x.view(u)
        top = u(1) << u(w - 1)
        return np.where((b >> u(w - 1)) != 0, ~b, b | top).astype(u)
    if x.dtype == np.int32:
        return x.view(np.uint32) ^ np.uint32(0x80000000)
    return x.astype(x.dtype, copy=False)


def bucket_of(keys: np.ndarray, st, lut: np.ndarray) -> np.ndarray:
    """The segment's bucket of every key (BucketFn::operator())."""
    KT = keys.dtype.type
    W = 8 * keys.dtype.itemsize
    kb, lmax = KT(int(st["kb"]) & ((1 << W) - 1)), KT(int(st["lmax"]) & ((1 << W) - 1))
    low = np.minimum(np.maximum(keys, kb) - kb, lmax).astype(np.uint64)
    pmask = np.uint64((1 << int(st["pwid"])) - 1)
    tmask = np.uint64((1 << int(st["twid"])) - 1)
    p = (low >> np.uint64(int(st["poff"]))) & pmask
    t = (((low >> np.uint64(int(st["rsh"]))) & tmask) << np.uint64(int(st["lsh"]))) & np.uint64(0xFFFFFFFF)
    e = lut[p.astype(np.int64)].astype(np.uint64)
    nbsh = (e >> np.uint64(16)) << np.uint64(8)
    j = ((t & np.uint64(0xFFFFFF)) * (nbsh & np.uint64(0xFFFFFF))) >> np.uint64(32)
    b = (e & np.uint64(0xFFFF)) + j
    if int(st["hvon"]):
        hv = KT(int(st["hv"]) & ((1 << W) - 1))
        b = b + (keys >= hv).astype(np.uint64) + (keys > hv).astype(np.uint64)
    return b.astype(np.int64)


def device_counts(totals: np.ndarray, gsums: np.ndarray, B: int) -> np.ndarray:
    """Bucket counts from the reduce pass's per-group exclusive prefixes and
    group sums (k_reduce)."""
    ng = (B + REDUCE_BINS - 1) // REDUCE_BINS
    t = totals[:ng * REDUCE_BINS].astype(np.int64).reshape(ng, REDUCE_BINS)
    nxt = np.concatenate([t[:, 1:], gsums[:ng, None].astype(np.int64)], axis=1)
    c = (nxt - t).reshape(-1)
    # bins past B inside the last group hold no prefix: their count is the
    # group sum minus the last valid prefix, all of it given to bin B - 1
    c = c[:B].copy()
    if B % REDUCE_BINS:
        g = ng - 1
        c[B - 1] = int(gsums[g]) - int(t[g, (B - 1) % REDUCE_BINS])
    return c

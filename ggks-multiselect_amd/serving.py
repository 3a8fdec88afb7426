"""Served request streams over resident vectors: workload, replay, checks.

A pool of vectors lives in HBM (one buffer per dtype); a seeded open-loop
trace of Poisson arrivals asks multiselect (1..256 uniform k) or top-K
(K log-uniform in [1, min(4096, len)]) questions about them; the library's
serving loop (ggks_server_*) runs the trace with continuous batching and
reports each request's latency.  Every answer is then checked against sorted
host copies of the pool (one lookup per output).  The reference has no
serving layer; per request the answer is the reference's sort-and-choose
(multiselectTimingFunctions.cu:41-95) or top-K (:155-177) on that vector.
"""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _lib, api, workloads as W

DTYPES = (np.float32, np.uint32, np.float64)  # drawn 6 : 3 : 1


def _unit(z):
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def make_pool(seed: int, npool: int, log2_lo: float = 12.0, log2_hi: float = 26.0) -> dict:
    """Pool spec: per vector its dtype, length, value family and seed, and its
    offset inside its dtype's resident buffer."""
    z = W.splitmix64(seed, 0, 4 * npool).reshape(npool, 4)
    lengths = np.floor(2.0 ** (log2_lo + _unit(z[:, 0]) * (log2_hi - log2_lo))).astype(np.uint64)
    r = z[:, 1] % np.uint64(10)
    dt = np.where(r < 6, 0, np.where(r < 9, 1, 2)).astype(np.int64)  # index into DTYPES
    fam = (z[:, 2] % np.uint64(3)).astype(np.int64)  # 0 uniform, 1 normal, 2 duplicate-heavy
    offsets = np.zeros(npool, np.uint64)
    totals = [0, 0, 0]
    for i in range(npool):
        offsets[i] = totals[dt[i]]
        totals[dt[i]] += int(lengths[i])
    return {"lengths": lengths, "dt": dt, "family": fam, "seeds": z[:, 3], "offsets": offsets,
            "totals": totals}


def upload_pool(pool: dict) -> dict:
    """Generate the pool on the device: {dtype: DeviceArray}."""
    bufs = {}
    for j, dt in enumerate(DTYPES):
        if pool["totals"][j]:
            bufs[dt] = api.DeviceArray(pool["totals"][j], dt)
    for i in range(pool["lengths"].size):
        dt = DTYPES[pool["dt"][i]]
        fam, n, sd = int(pool["family"][i]), int(pool["lengths"][i]), int(pool["seeds"][i])
        off = int(pool["offsets"][i])
        if dt == np.uint32:
            lo, hi = (0, 1023) if fam == 2 else (0, 2**32 - 1)
            api.generate(dt, "uniform", sd, n, lo=lo, hi=hi, out=bufs[dt], offset=off)
        elif fam == 2:
            api.generate(dt, "levels", sd, n, lo=0, hi=1023, out=bufs[dt], offset=off)
        else:
            api.generate(dt, "normal" if fam == 1 else "uniform", sd, n, out=bufs[dt], offset=off)
    return bufs


def sorted_copies(pool: dict, bufs: dict, threads: int = 16) -> dict:
    """Host copies of the resident buffers with every vector sorted ascending
    (NaN-free and without -0.0, so numpy's order is the radix order)."""
    host = {dt: b.to_host() for dt, b in bufs.items()}

    def sort_one(i):
        dt = DTYPES[pool["dt"][i]]
        o, n = int(pool["offsets"][i]), int(pool["lengths"][i])
        host[dt][o:o + n].sort()

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(sort_one, range(pool["lengths"].size)))
    return host


def make_trace(seed: int, pool: dict, rate: float, seconds: float, max_k: int = 256,
               max_topk: int = 4096) -> dict:
    """Poisson arrivals at `rate` per second for `seconds`; host k and output
    buffers and the ggks_request records pointing into them."""
    n_est = int(rate * seconds * 1.2 + 100)
    gaps = -np.log1p(-_unit(W.splitmix64(seed, 0, n_est))) / rate
    arrival = np.cumsum(gaps)
    while arrival[-1] < seconds:  # extend in the unlikely case the estimate was short
        more = -np.log1p(-_unit(W.splitmix64(seed, arrival.size, n_est))) / rate
        arrival = np.concatenate([arrival, arrival[-1] + np.cumsum(more)])
    arrival = arrival[arrival < seconds]
    n = arrival.size
    z = W.splitmix64(seed ^ 0x5EED, 0, 4 * n).reshape(n, 4) if n else np.zeros((0, 4), np.uint64)
    vec = (z[:, 0] % np.uint64(pool["lengths"].size)).astype(np.int64)
    length = pool["lengths"][vec]
    topk = (z[:, 1] & np.uint64(1)).astype(bool)
    nk = (1 + z[:, 2] % np.uint64(max_k)).astype(np.uint32)
    kmax = np.minimum(length, max_topk).astype(np.float64)
    K = np.clip(np.floor(2.0 ** (_unit(z[:, 3]) * np.log2(kmax))), 1, kmax).astype(np.uint32)
    num_k = np.where(topk, K, nk).astype(np.uint32)
    # k lists of the multiselect requests, back to back
    kcount = np.where(topk, 0, nk).astype(np.int64)
    koff = np.concatenate([[0], np.cumsum(kcount)[:-1]]).astype(np.int64) if n else kcount
    total_k = int(kcount.sum())
    owner = np.repeat(np.arange(n), kcount)
    kz = W.splitmix64(seed ^ 0xC0FFEE, 0, total_k) if total_k else np.zeros(0, np.uint64)
    kvals = (kz % length[owner] + np.uint64(1)).astype(np.uint32)
    # outputs per dtype, back to back in arrival order
    dt = pool["dt"][vec]
    out_off = np.zeros(n, np.int64)
    outs = []
    for j, d in enumerate(DTYPES):
        m = dt == j
        c = num_k[m].astype(np.int64)
        out_off[m] = np.concatenate([[0], np.cumsum(c)[:-1]]) if c.size else c
        outs.append(np.zeros(int(c.sum()), d))
    reqs = np.zeros(n, _lib.REQUEST_DTYPE)
    reqs["offset"] = pool["offsets"][vec]
    reqs["length"] = length
    reqs["h_k"] = np.where(topk, 0, kvals.ctypes.data + 4 * koff).astype(np.uint64)
    eb = np.array([np.dtype(d).itemsize for d in DTYPES])
    base = np.array([o.ctypes.data for o in outs], np.uint64)
    reqs["h_out"] = base[dt] + (out_off * eb[dt]).astype(np.uint64)
    reqs["tag"] = np.arange(n, dtype=np.uint64)
    reqs["num_k"] = num_k
    reqs["kind"] = np.where(topk, _lib.KIND_TOPK, _lib.KIND_BUCKET)
    reqs["dtype"] = np.array([api.dtype_code(d) for d in DTYPES])[dt]
    return {"arrival": arrival, "reqs": reqs, "vec": vec, "topk": topk, "num_k": num_k,
            "kvals": kvals, "koff": koff, "dt": dt, "out_off": out_off, "outs": outs,
            "length": length}


def verify(pool: dict, host_sorted: dict, tr: dict) -> int:
    """Count of requests whose outputs differ from the sorted copies."""
    bad = np.zeros(tr["vec"].size, bool)
    for j, d in enumerate(DTYPES):
        m = np.nonzero(tr["dt"] == j)[0]
        if m.size == 0:
            continue
        nk = tr["num_k"][m].astype(np.int64)
        rid = np.repeat(m, nk)                                   # request of each output
        pos = np.arange(int(nk.sum())) - np.repeat(np.cumsum(nk) - nk, nk)  # j within request
        n = tr["length"][rid].astype(np.int64)
        top = tr["topk"][rid]
        kk = np.where(top, 0, tr["kvals"][np.minimum(tr["koff"][rid] + pos,
                                                     max(tr["kvals"].size - 1, 0))]
                      if tr["kvals"].size else 0).astype(np.int64)
        idx = np.where(top, n - 1 - pos, n - kk)                 # ascending index
        want = host_sorted[d][pool["offsets"][tr["vec"][rid]].astype(np.int64) + idx]
        got = tr["outs"][j][tr["out_off"][rid] + pos]
        u = np.uint64 if np.dtype(d).itemsize == 8 else np.uint32
        wrong = want.view(u) != got.view(u)
        bad[np.unique(rid[wrong])] = True
    return int(bad.sum())


def run(server: api.Server, pool: dict, host_sorted: dict, rate: float, seconds: float,
        seed: int) -> dict:
    """Replay one trace and summarise it."""
    tr = make_trace(seed, pool, rate, seconds)
    s0 = server.stats()
    import time
    t0 = time.perf_counter()
    lat = server.run_trace(tr["reqs"], tr["arrival"])
    wall = time.perf_counter() - t0
    s1 = server.stats()
    ok = lat >= 0
    eb = np.array([np.dtype(d).itemsize for d in DTYPES])[tr["dt"]]
    elems = float(tr["length"][ok].sum())
    nbad = verify(pool, host_sorted, tr)
    return {
        "rate_per_s": rate, "seconds": seconds, "requests": int(tr["arrival"].size),
        "completed": int(ok.sum()), "refused": int((~ok).sum()),
        "p50_ms": round(float(np.percentile(lat[ok], 50)) * 1e3, 3) if ok.any() else None,
        "p99_ms": round(float(np.percentile(lat[ok], 99)) * 1e3, 3) if ok.any() else None,
        "mean_ms": round(float(lat[ok].mean()) * 1e3, 3) if ok.any() else None,
        "gelem_per_s": round(elems / wall / 1e9, 2),
        "offered_gelem_per_s": round(float(tr["length"].sum()) / seconds / 1e9, 2),
        "hbm_frac": round(2.0 * float((tr["length"][ok] * eb[ok]).sum()) / wall / 8e12, 4),
        "wall_s": round(wall, 2),
        "batches": s1["batches"] - s0["batches"],
        "mean_batch": round((s1["completed"] - s0["completed"]) /
                            max(1, s1["batches"] - s0["batches"]), 1),
        "max_pending": s1["max_pending"], "peak_scratch_gb": round(s1["peak_scratch_bytes"] / 1e9, 2),
        "bit_exact": nbad == 0, "mismatched_requests": nbad,
    }

"""Autotuner of the batched path's band table.

For every (dtype, floor(log2 length)) band it times one representative batch
(segments with lengths uniform in [2^b, 2^(b+1)), about `elems` elements in
all, each asking 16 uniform k values or a top-K of 64) under each candidate
(bucket budget, chunk size, direct route) and keeps the fastest.  Timing is
the mean of graph replays on the device; every candidate's outputs must be
bit-identical to the first candidate's, so a faster but wrong choice cannot be
taken.  The result is a JSON table that ``apply()`` loads into the library
(ggks_tuning_set).  The reference fixes its bucket count and block shape
(bucketMultiselect.cu:841, :868, multiselectTimingFunctions.cu:105-113) and
leaves its size cutoffs commented out (bucketMultiselect.cu:849-866); this
measures them per band instead.

Usage: python ggks-multiselect_amd/autotune.py [--dtypes f32,f64] [--bands 10-26]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PATH = os.path.join(HERE, "tuning_gfx950.json")
DTYPES = {"f32": np.float32, "f64": np.float64, "u32": np.uint32, "i32": np.int32}
LDS_KEYS_BYTES = 65536  # the finisher keeps a whole segment in LDS up to this size


def _pkg():
    try:
        import ggks_multiselect_amd as pkg  # registered by __graft_entry__.load_package
        return pkg
    except ImportError:
        sys.path.insert(0, os.path.dirname(HERE))
        from __graft_entry__ import load_package
        return load_package()


def band_batch(pkg, dtype, band: int, elems: int, seed: int):
    """Device data, segment table and k list of one band's probe batch."""
    W = pkg.workloads
    lo = 1 << band
    nseg = int(max(8, min(4096, elems // int(1.5 * lo))))
    z = W.splitmix64(seed + band, 0, 2 * nseg)
    lengths = (lo + z[:nseg] % np.uint64(lo)).astype(np.uint64)
    topk = (z[nseg:] & np.uint64(3)) == 0  # a quarter top-K
    num_k = np.where(topk, np.minimum(64, lengths), 16).astype(np.uint32)
    ks = [(W.splitmix64(int(seed + i), 0, 16) % lengths[i] + np.uint64(1)).astype(np.uint32)
          for i in np.nonzero(~topk)[0]]
    k = np.concatenate(ks) if ks else np.zeros(0, np.uint32)
    table = pkg.segment_table(lengths, np.where(topk, 2, 0), num_k)
    d = pkg.DeviceArray(int(lengths.sum()), dtype)
    fam = "uniform" if np.dtype(dtype).kind in "ui" else "normal"
    hi = 0xFFFFFFFF if dtype == np.uint32 else 2**31 - 1
    lo_v = 0 if dtype == np.uint32 else -2**31
    pkg.generate(dtype, fam, seed, d.n, lo=lo_v, hi=hi, out=d)
    return d, table, k


def candidates(dtype, band: int):
    cap = LDS_KEYS_BYTES // np.dtype(dtype).itemsize
    out = []
    if (1 << (band + 1)) - 1 <= cap:
        out.append((10, 16, True))
    for bl in (8, 10, 11, 12, 13, 14, 15):
        if bl > band - 2:
            continue
        for cl in (13, 14, 15, 16, 17, 18, 19, 20):
            if cl > band + 1 or cl < band - 9:  # 1 .. 512 chunks per segment
                continue
            out.append((bl, cl, False))
    return out


def time_batch(pkg, dtype, d, table, k, reps: int):
    b = pkg.Batch(dtype, table, k)
    out = pkg.DeviceArray(max(b.num_outputs, 1), dtype)
    b.run(d, out, graph=True)
    b.sync()
    res = out.to_host()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.run(d, out, graph=True)
    b.sync()
    dt = (time.perf_counter() - t0) / reps
    b.close()
    return dt, res


def tune(dtypes=("f32", "f64"), bands=range(10, 27), elems: int = 1 << 26, reps: int = 5,
         seed: int = 77, log=print) -> dict:
    pkg = _pkg()
    table_out = {}
    for name in dtypes:
        dt = DTYPES[name]
        table_out[name] = {}
        for band in bands:
            d, table, k = band_batch(pkg, dt, band, elems, seed)
            best, ref = None, None
            try:
                for bl, cl, direct in candidates(dt, band):
                    pkg.tuning_set(dt, band, bl if not direct else 10, cl if not direct else 16,
                                   direct)
                    t, res = time_batch(pkg, dt, d, table, k, reps)
                    if ref is None:
                        ref = res
                    elif not np.array_equal(res.view(np.uint8), ref.view(np.uint8)):
                        raise RuntimeError(f"{name} band {band} ({bl},{cl},{direct}): outputs differ")
                    if best is None or t < best[0]:
                        best = (t, bl, cl, direct)
            finally:
                pkg.tuning_reset()
            t, bl, cl, direct = best
            table_out[name][str(band)] = [bl, cl, bool(direct)]
            log(f"{name} band 2^{band}: buckets 2^{bl} chunk 2^{cl} direct={direct} "
                f"{d.n / t / 1e9:.1f} Gelem/s", flush=True)
            d.free()
    return table_out


def apply(path: str = DEFAULT_PATH) -> bool:
    """Load a tuned table into the library; False if there is none."""
    if not os.path.exists(path):
        return False
    pkg = _pkg()
    with open(path) as f:
        doc = json.load(f)
    for name, bands in doc.get("bands", {}).items():
        for band, (bl, cl, direct) in bands.items():
            pkg.tuning_set(DTYPES[name], int(band), int(bl), int(cl), bool(direct))
            # the same key width shares a table when only one was tuned
            if name == "f32":
                for other in ("u32", "i32"):
                    if other not in doc["bands"]:
                        pkg.tuning_set(DTYPES[other], int(band), int(bl), int(cl), bool(direct))
    return True


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--dtypes", default="f32,f64")
    ap.add_argument("--bands", default="10-26")
    ap.add_argument("--elems", type=int, default=1 << 26)
    ap.add_argument("--out", default=DEFAULT_PATH)
    a = ap.parse_args()
    b0, b1 = (int(v) for v in a.bands.split("-"))
    t0 = time.time()
    res = tune(a.dtypes.split(","), range(b0, b1 + 1), a.elems)
    doc = {"device": "gfx950", "method": "fastest bit-identical candidate per band, mean of "
           "graph replays", "elems_per_band": a.elems, "bands": res}
    with open(a.out, "w") as f:
        json.dump(doc, f, indent=1)
    print(f"wrote {a.out} in {time.time() - t0:.0f}s", flush=True)


if __name__ == "__main__":
    main()

"""Autotuner of the band tables: the batched path's, and the single-vector
one.

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

Two more modes:
- single: per (dtype, floor(log2 n)) the bucket budget and chunk count of
  the single-vector entry points (one vector of 2^b keys, 1024 uniform k),
  timed as back-to-back stream-ordered selects, fastest bit-identical
  candidate kept (ggks_tuning_single_set);
- mix: a coordinate descent of the batch bands inside the whole mixed batch
  (the 4096-vector workload), because a band probed alone pays for its
  direct segments' finishers, which in the mix run beside the pipeline; a
  band's entry changes only when the whole batch gets faster by more than
  the replay noise.

Usage: python ggks-multiselect_amd/autotune.py [--dtypes f32,f64] [--bands 10-26]
       python ggks-multiselect_amd/autotune.py --mode single [--bands 16-28]
       python ggks-multiselect_amd/autotune.py --mode mix
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


def single_candidates(band: int):
    out = []
    for bl in (11, 12, 13, 14, 15):
        if bl > band - 4:
            continue
        for ch in (64, 128, 256, 384, 512):
            if ch * 4096 > (1 << band):
                continue
            out.append((bl, ch))
    return out


def tune_single(dtypes=("f32", "f64", "u32", "i32"), bands=range(16, 29), reps: int = 10,
                numk: int = 1024, seed: int = 91, log=print) -> dict:
    """Per (dtype, band): the fastest bit-identical (bucket budget, chunks)
    of the single-vector select on one resident vector of 2^band keys."""
    pkg = _pkg()
    lib = pkg.api.lib()
    W = pkg.workloads
    out = {}
    for name in dtypes:
        dt = DTYPES[name]
        out[name] = {}
        for band in bands:
            n = 1 << band
            if np.dtype(dt).itemsize == 8 and band > 27:
                continue
            fam = "normal" if np.dtype(dt).kind == "f" else "uniform"
            hi = 0xFFFFFFFF if dt == np.uint32 else 2**31 - 1
            lo_v = 0 if dt == np.uint32 else -2**31
            d = pkg.generate(dt, fam, seed + band, n, lo=lo_v, hi=hi)
            k = W.k_uniform(seed, n, min(numk, n // 4))
            dk = pkg.DeviceArray.from_host(k)
            dout = pkg.DeviceArray(k.size, dt)
            ctx = pkg.Context(0)
            best, ref = None, None
            for bl, ch in single_candidates(band):
                for _ in range(2):
                    pkg.multiselect_async(d, dk, dout, ctx, buckets_log2=bl, blocks=ch)
                pkg.api.check(lib.ggks_sync())
                res = dout.to_host()
                t0 = time.perf_counter()
                for _ in range(reps):
                    pkg.multiselect_async(d, dk, dout, ctx, buckets_log2=bl, blocks=ch)
                pkg.api.check(lib.ggks_sync())
                t = (time.perf_counter() - t0) / reps
                if ref is None:
                    ref = res
                elif not np.array_equal(res.view(np.uint8), ref.view(np.uint8)):
                    raise RuntimeError(f"{name} 2^{band} ({bl},{ch}): outputs differ")
                if best is None or t < best[0]:
                    best = (t, bl, ch)
            if best is None:
                continue
            t, bl, ch = best
            out[name][str(band)] = [bl, ch]
            log(f"single {name} 2^{band}: buckets 2^{bl}, {ch} chunks, {t * 1e6:.1f} us, "
                f"{n / t / 1e9:.1f} Gelem/s", flush=True)
            ctx.close()
            d.free()
    return out


def tune_mix(dtype="f32", nseg: int = 4096, reps: int = 8, rounds: int = 1, margin: float = 0.02,
             path: str = DEFAULT_PATH, log=print) -> dict:
    """Coordinate descent of the batch bands inside the mixed batch: for each
    band of the workload, every candidate entry is tried with the rest of the
    table fixed; the best is kept if the whole batch's replay is faster than
    the current entry's by more than `margin`."""
    pkg = _pkg()
    dt = DTYPES[dtype]
    apply(path)
    spec = pkg.workloads.batch_spec(4096, nseg, 10, 22)
    table = pkg.segment_table(spec["lengths"], spec["kinds"], spec["num_k"])
    d = pkg.DeviceArray(int(spec["lengths"].sum()), dt)
    for i, srow in enumerate(table):
        pkg.generate(dt, "normal" if spec["normal"][i] else "uniform", int(spec["seeds"][i]),
                     int(srow["length"]), out=d, offset=int(srow["offset"]))
    bands = sorted({int(np.log2(float(n))) for n in spec["lengths"]})

    def run():
        t, res = time_batch(pkg, dt, d, table, spec["k"], reps)
        return t, res

    base_t, ref = run()
    log(f"mix: start {base_t * 1e3:.3f} ms per replay", flush=True)
    changed = {}
    for _ in range(rounds):
        for band in bands:
            cur = pkg.tuning_get(dt, band)
            best = (base_t, cur)
            for bl, cl, direct in candidates(dt, band):
                cand = (bl if not direct else 10, cl if not direct else 16, direct)
                if cand == cur:
                    continue
                pkg.tuning_set(dt, band, *cand)
                t, res = run()
                if not np.array_equal(res.view(np.uint8), ref.view(np.uint8)):
                    raise RuntimeError(f"mix band {band} {cand}: outputs differ")
                if t < best[0]:
                    best = (t, cand)
            if best[1] != cur and best[0] < base_t * (1 - margin):
                pkg.tuning_set(dt, band, *best[1])
                base_t = best[0]
                changed[str(band)] = list(best[1])
                log(f"mix: band 2^{band} -> {best[1]}: {base_t * 1e3:.3f} ms", flush=True)
            else:
                pkg.tuning_set(dt, band, *cur)
    table_now = {str(b): list(pkg.tuning_get(dt, b)) for b in bands}
    d.free()
    return {"ms_per_replay": base_t * 1e3, "changed": changed, "bands": table_now}


def apply(path: str = DEFAULT_PATH) -> bool:
    """Load a tuned table into the library; False if there is none."""
    if not os.path.exists(path):
        return False
    pkg = _pkg()
    with open(path) as f:
        doc = json.load(f)
    for name, bands in doc.get("single", {}).items():
        for band, (bl, ch) in bands.items():
            pkg.tuning_single_set(DTYPES[name], int(band), int(bl), int(ch))
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
    ap.add_argument("--mode", default="bands", choices=["bands", "single", "mix"])
    ap.add_argument("--dtypes", default=None)
    ap.add_argument("--bands", default=None)
    ap.add_argument("--elems", type=int, default=1 << 26)
    ap.add_argument("--out", default=DEFAULT_PATH)
    a = ap.parse_args()
    t0 = time.time()
    doc = {}
    if os.path.exists(a.out):
        with open(a.out) as f:
            doc = json.load(f)
    doc["device"] = "gfx950"
    if a.mode == "bands":
        b0, b1 = (int(v) for v in (a.bands or "10-26").split("-"))
        doc["bands"] = tune((a.dtypes or "f32,f64").split(","), range(b0, b1 + 1), a.elems)
        doc["method"] = "fastest bit-identical candidate per band, mean of graph replays"
        doc["elems_per_band"] = a.elems
    elif a.mode == "single":
        b0, b1 = (int(v) for v in (a.bands or "16-28").split("-"))
        doc["single"] = tune_single((a.dtypes or "f32,f64,u32,i32").split(","), range(b0, b1 + 1))
        doc["single_method"] = ("fastest bit-identical (buckets, chunks) per dtype and band, one "
                                "resident vector of 2^band keys, 1024 uniform k, mean of "
                                "back-to-back selects")
    else:
        r = tune_mix((a.dtypes or "f32"), path=a.out)
        # the mixed batch spans bands 10..21: the other bands keep their entries
        doc.setdefault("bands", {}).setdefault(a.dtypes or "f32", {}).update(r["bands"])
        doc["mix"] = {"ms_per_replay": round(r["ms_per_replay"], 4), "changed": r["changed"],
                      "method": "coordinate descent of each band's entry inside the 4096-vector "
                                "mixed batch, kept when the whole replay is > 2 % faster"}
    with open(a.out, "w") as f:
        json.dump(doc, f, indent=1)
    print(f"wrote {a.out} in {time.time() - t0:.0f}s", flush=True)


if __name__ == "__main__":
    main()

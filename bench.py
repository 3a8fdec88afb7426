#!/usr/bin/env python3
"""Benchmark of the selection hot path on one MI355X.

Default workload: bucket multiselect over n = 2^28 float32 N(0,1) values
(seed 1) with 1024 k values uniform in [1, n] (seed 101, sorted and
deduplicated), the vector resident in HBM.  A step is one select through the
stream-ordered C-ABI entry point (device k list, device outputs); the timed
region is K back-to-back steps between two stream synchronisations, and HIP
events recorded by the library on its stream time every kernel stage of every
step.

Prints ONE JSON line.  value = n / time per step (Gelem/s).  roofline is for
the dominant kernel (the larger of the two streaming passes, each of which
reads n * 4 bytes): achieved = n * 4 B / its mean event duration.
cpu_baseline = the reference's sort-and-choose (its vendored Thrust 1.4, OpenMP
backend, oracle/_ref) on the host cores, over the first 2^24 elements of the
same vector with 1024 k values, timed for a bounded ~10 s.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
from __graft_entry__ import load_package  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "Gelem/s and %HBM peak, batched vs plain (bit-exact); p50/p99 request latency"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--dtype", default="f32", choices=["f32", "f64", "u32", "i32"])
    ap.add_argument("--numk", type=int, default=1024)
    ap.add_argument("--kind", default="bucket", choices=["bucket", "radix"])
    ap.add_argument("--blocks", type=int, default=0, help="chunk count (0 = tuned)")
    ap.add_argument("--buckets-log2", type=int, default=0, help="bucket budget (0 = tuned)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the timed steps, copy the vector once with the library's plain "
                         "16-byte copy kernel (known-size reference for memory-traffic counters)")
    return ap.parse_args()


def make_input(pkg, dtype: str, n: int):
    W = pkg.workloads
    if dtype == "f32":
        return W.normal_f32(1, n)
    if dtype == "f64":
        return W.normal_f64(2, n)
    if dtype == "u32":
        return W.uint_range(3, n, 0, 0xFFFFFFFF)
    return W.int_range(3, n, -(2 ** 31), 2 ** 31 - 1)


def cpu_baseline(pkg, x: np.ndarray, numk: int, budget_s: float):
    """Reference sort-and-choose on host cores over a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # checker / baseline only
    ns = min(x.size, 1 << 24)
    sample = np.ascontiguousarray(x[:ns])
    ks = pkg.workloads.k_uniform(101, ns, numk)
    threads = os.cpu_count() or 1
    kind = "reference" if O.ref_available() else "port"
    times, res = [], None
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        if kind == "reference":
            res, _sort_ms = O.ref_sort_choose(sample, ks, threads=threads)
        else:
            res = O.parallel_sort_choose(sample, ks, threads=threads)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start >= budget_s or len(times) >= 20:
            break
    t = float(np.median(times))
    # the GPU path on the same sample must return the same values
    d = pkg.DeviceArray.from_host(sample)
    gpu = pkg.multiselect(d, ks)
    u = np.uint64 if sample.dtype == np.float64 else np.uint32
    parity = bool(np.array_equal(gpu.view(u), res.view(u)))
    return {
        "value": round(ns / t / 1e9, 4),
        "unit": "Gelem/s",
        "cores": threads,
        "kind": kind,
        "sample": (f"first 2^{int(np.log2(ns))} elements of the workload vector, {ks.size} "
                   f"uniform k; {'vendored Thrust 1.4 OMP sort' if kind == 'reference' else 'libstdc++ parallel sort'}"
                   f" + gather at n-k, median of {len(times)} runs (copy-in included)"),
        "ms": round(t * 1e3, 2),
        "gpu_parity": parity,
    }


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported "
              "(multi-GPU sharding is described in DESIGN.md, not built)", file=sys.stderr)
        sys.exit(2)
    pkg = load_package()
    from ggks_multiselect_amd import _lib as L  # noqa: E402
    n = 1 << a.log2n
    t_gen = time.perf_counter()
    x = make_input(pkg, a.dtype, n)
    k = pkg.workloads.k_uniform(101, n, a.numk)
    t_gen = time.perf_counter() - t_gen
    esz = x.dtype.itemsize
    print(f"# generated n=2^{a.log2n} {a.dtype} and {k.size} k in {t_gen:.1f}s", file=sys.stderr,
          flush=True)

    d = pkg.DeviceArray.from_host(x)
    dk = pkg.DeviceArray.from_host(k)
    dout = pkg.DeviceArray(k.size, x.dtype)
    ctx = pkg.Context(0)
    lib = L.lib()
    L.check(lib.ggks_ctx_reserve(ctx.handle, L.KIND_BUCKET if a.kind == "bucket" else L.KIND_RADIX,
                                 pkg.api.dtype_code(x.dtype), n, k.size))
    # correctness of the timed configuration against the sync path's answer
    # on the same data is covered by the tests; here the outputs of the first
    # and last timed step must agree
    opt = {"blocks": a.blocks, "buckets_log2": a.buckets_log2}
    for _ in range(a.warmup):
        pkg.multiselect_async(d, dk, dout, ctx, kind=a.kind, **opt)
    L.check(lib.ggks_sync())
    first = dout.to_host()
    per = L.NSTAGES + 1
    L.check(lib.ggks_ctx_profile_arm(ctx.handle, a.steps))
    L.check(lib.ggks_sync())
    t0 = time.perf_counter()
    for _ in range(a.steps):
        pkg.multiselect_async(d, dk, dout, ctx, kind=a.kind, **opt)
    L.check(lib.ggks_sync())
    wall = time.perf_counter() - t0
    ms = np.zeros(a.steps * per, np.float32)
    L.check(lib.ggks_ctx_profile_read(ctx.handle, ms.ctypes.data, a.steps))
    ms = ms.reshape(a.steps, per)
    last = dout.to_host()
    consistent = bool(np.array_equal(first.view(np.uint8), last.view(np.uint8)))
    L.check(lib.ggks_ctx_profile_arm(ctx.handle, 0))
    if a.copy_probe:
        scratch = pkg.DeviceArray(n, x.dtype)
        L.check(lib.ggks_d2d(scratch.ptr, d.ptr, n * esz))
        del scratch

    step_ms = float(ms[:, L.NSTAGES].mean())
    wall_ms = wall * 1e3 / a.steps
    stage = {name: round(float(ms[:, i].mean()) * 1e3, 2) for i, name in enumerate(L.STAGES)}
    dom = "count" if stage["count"] >= stage["compact"] else "compact"
    dom_us = stage[dom]
    achieved = n * esz / (dom_us * 1e-6) / 1e9
    call_frac = 2 * n * esz / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
    value = n / (step_ms * 1e-3) / 1e9

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tt = json.load(f)
        key = f"{a.kind}_{a.dtype}_2^{a.log2n}_{dom}"
        traffic = tt.get(key)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Gelem/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(step_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": a.dtype,
        "data": "synthetic (seeded SplitMix64 generators), resident in HBM",
        "config": {
            "workload": f"{a.kind} multiselect, n=2^{a.log2n} {a.dtype} "
                        f"{'N(0,1)' if a.dtype in ('f32', 'f64') else 'uniform'}, "
                        f"{k.size} uniform k (sorted, deduplicated)",
            "n": n, "num_k": int(k.size), "blocks": a.blocks, "buckets_log2": a.buckets_log2,
            "hbm_frac_call": round(call_frac, 4),
            "alg_bytes_per_step": 2 * n * esz,
            "wall_ms_per_step": round(wall_ms, 4),
            "stage_us": stage,
            "outputs_stable": consistent,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_count" if dom == "count" else "k_compact",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(pkg, x, a.numk, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

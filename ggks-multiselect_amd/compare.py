"""Algorithm comparison, crossover search and result summaries in the
reference's file formats.

- ``compare``: for one problem (size, k list, value family, k family) run
  each algorithm ``tests`` times in a shuffled order on the same device
  vector, check every output against sort-and-choose, and append one CSV row
  per test in the reference's layout (compareMultiselect.hpp:159-187)::

      size,numKs,arrayDistribution,kDistribution,name,ms,name,ms,...,errors

- ``crossover``: the number of k values at which bucket multiselect stops
  beating sort-and-choose for a vector size, found as the reference does
  (analyzeMultiselect.cu:209-274): grow numK while bucket multiselect wins
  (x2 below 5000, x1.5 below 10000, x1.2 above), then bisect the bracket.
- ``summarize``: the reference's summaries of such rows
  (readMultiselectOutputfile.cpp:209-345): per (size, numK) and per size,
  each algorithm's mean / max / min time and wins (Summary, Overview) and
  its time over the first algorithm's, mean / max / min (RatioSummary,
  RatioOverview).

The algorithms are the library's: "Sort and Choose Multiselect" (the
reference's baseline, multiselectTimingFunctions.cu:41-95, as a device radix
sort plus picks), "Bucket Multiselect" and "Radix Multiselect".  Times are
wall-clock milliseconds of the synchronous C-ABI call on a device-resident
vector, host k list in and host outputs back, as the reference's timing
functions measure (k upload and output download included).

Usage (from the repository root)::

    python ggks-multiselect_amd/compare.py crossover --sizes 20,22,24 --csv out.csv
    python ggks-multiselect_amd/compare.py summarize out
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import time
from collections import OrderedDict

import numpy as np

ALGORITHMS = OrderedDict([
    ("Sort and Choose Multiselect", "sort_choose"),
    ("Bucket Multiselect", "bucket"),
    ("Radix Multiselect", "radix"),
])
K_FAMILIES = ("uniform random", "uniform", "normal random", "cluster", "sectioned")


def k_list(W, family: str, seed: int, count: int, n: int) -> np.ndarray:
    """count k values of the reference's k families (generateProblems.cu:
    489-584), from the library's seeded generators."""
    if family == "uniform random":
        return W.k_uniform(seed, n, count, dedup=False)
    if family == "uniform":
        return W.k_even(count, n)
    if family == "normal random":
        return W.k_normal(seed, count, n)
    if family == "cluster":
        return W.k_cluster9(seed, count, n)
    if family == "sectioned":
        return W.k_sectioned(seed, count, n)
    raise ValueError(f"unknown k family {family!r}")


def _run(pkg, alg: str, d, k: np.ndarray) -> tuple:
    """(values, milliseconds) of one synchronous call."""
    lib = pkg.api.lib()
    out = np.empty(k.size, d.dtype)
    kk = np.ascontiguousarray(k, np.uint32)
    code = pkg.api.dtype_code(d.dtype)
    t0 = time.perf_counter()
    if alg == "sort_choose":
        rc = lib.ggks_sort_choose(None, code, d.ptr, len(d), kk.ctypes.data, kk.size,
                                  out.ctypes.data)
        pkg.api.check(rc)
    else:
        out = pkg.multiselect(d, kk, kind=alg)
    return out, (time.perf_counter() - t0) * 1e3


def compare(pkg, n: int, num_k: int, dtype=np.float32, family: str = "uniform",
            kfamily: str = "uniform random", tests: int = 3, seed: int = 1,
            algorithms=("Sort and Choose Multiselect", "Bucket Multiselect"),
            csv_path: str | None = None) -> dict:
    """Time each algorithm `tests` times (shuffled order per test, the same
    vector and k list), check outputs against the first algorithm, append
    the rows to csv_path; returns mean times, wins and the error count."""
    W = pkg.workloads
    rng = np.random.default_rng(seed)
    times = {a: [] for a in algorithms}
    wins = {a: 0 for a in algorithms}
    errors = 0
    rows = []
    for t in range(tests):
        d = pkg.generate(dtype, family, seed + t, n, hi=n if family == "sorted" else 0)
        k = k_list(W, kfamily, seed + 1000 + t, num_k, n)
        res, ms = {}, {}
        if t == 0:  # untimed: first-call costs (scratch growth, library set-up)
            for a in algorithms:
                _run(pkg, ALGORITHMS[a], d, k)
        for j in rng.permutation(len(algorithms)):
            a = algorithms[j]
            res[a], ms[a] = _run(pkg, ALGORITHMS[a], d, k)
        ref = res[algorithms[0]]
        u = np.uint64 if np.dtype(dtype).itemsize == 8 else np.uint32
        flag = sum(int(np.count_nonzero(res[a].view(u) != ref.view(u))) for a in algorithms[1:])
        errors += flag
        for a in algorithms:
            times[a].append(ms[a])
        wins[min(algorithms, key=lambda a: ms[a])] += 1
        row = [n, num_k, family, kfamily]
        for a in algorithms:
            row += [a, round(ms[a], 4)]
        rows.append(row + [flag])
        d.free()
    if csv_path:
        with open(csv_path, "a", newline="") as f:
            csv.writer(f).writerows(rows)
    return {"n": n, "num_k": num_k, "mean_ms": {a: float(np.mean(v)) for a, v in times.items()},
            "wins": wins, "errors": errors}


def crossover(pkg, n: int, dtype=np.float32, family: str = "uniform",
              kfamily: str = "uniform random", tests: int = 3, start: int = 10,
              max_k: int | None = None, csv_path: str | None = None, log=None) -> dict:
    """Smallest numK (to within 1) at which bucket multiselect stops beating
    sort-and-choose on more tests than it loses, by the reference's bracket
    and bisection.  None when bucket multiselect still wins at max_k."""
    algs = ("Sort and Choose Multiselect", "Bucket Multiselect")
    max_k = max_k or n
    probes = []

    def bucket_wins(c):
        r = compare(pkg, n, c, dtype, family, kfamily, tests, seed=17 + len(probes),
                    algorithms=algs, csv_path=csv_path)
        probes.append(r)
        if log:
            log(f"n={n} numK={c}: {r['mean_ms']} wins={r['wins']}")
        return r["wins"][algs[1]] > r["wins"][algs[0]]

    i = a = start
    while bucket_wins(i):
        a = i
        if i >= max_k:
            return {"n": n, "crossover_num_k": None, "bracket": [a, None], "probes": probes}
        i = int(i * 2) if i < 5000 else int(i * 1.5) if i < 10000 else int(i * 1.2)
        i = min(i, max_k)
    b = i
    if b == start:
        return {"n": n, "crossover_num_k": start, "bracket": [None, b], "probes": probes}
    while (b - a) // 2 > 1:
        c = (a + b) // 2
        if bucket_wins(c):
            a = c
        else:
            b = c
    return {"n": n, "crossover_num_k": (a + b) // 2, "bracket": [a, b], "probes": probes}


# ------------------------------------------------------------- summaries
def read_rows(paths) -> list:
    """Parse result rows: (size, numK, arrayDistribution, kDistribution,
    [(name, ms), ...], errors).  Blank and non-row lines are skipped."""
    rows = []
    for p in paths:
        with open(p) as f:
            for rec in csv.reader(f):
                rec = [x for x in rec if x != ""]
                if len(rec) < 7 or not rec[0].strip().isdigit():
                    continue
                algs = [(rec[i], float(rec[i + 1])) for i in range(4, len(rec) - 1, 2)]
                rows.append((int(rec[0]), int(rec[1]), rec[2], rec[3], algs, int(rec[-1])))
    return rows


def _stats(rows) -> tuple:
    names = [a for a, _ in rows[0][4]]
    t = np.array([[ms for _, ms in r[4]] for r in rows], np.float64)
    wins = np.bincount(t.argmin(axis=1), minlength=len(names))
    ratio = t / t[:, :1]
    errors = int(sum(r[5] for r in rows))
    return names, t, wins, ratio, errors


def summarize(base: str, outdir: str | None = None) -> dict:
    """Write Summary<base>.csv, Overview<base>.csv, RatioSummary<base>.csv
    and RatioOverview<base>.csv from every file whose name starts with
    <base> (readMultiselectOutputfile.cpp:209-345 layouts); returns their
    paths."""
    d, b = os.path.split(base)
    d = d or "."
    paths = sorted(p for p in glob.glob(os.path.join(d, b + "*")) if os.path.isfile(p))
    rows = read_rows(paths)
    if not rows:
        raise ValueError(f"no result rows in files starting with {base!r}")
    outdir = outdir or d
    out = {k: os.path.join(outdir, f"{k}{b}.csv")
           for k in ("Summary", "Overview", "RatioSummary", "RatioOverview")}
    by_nk = OrderedDict()
    by_n = OrderedDict()
    for r in sorted(rows, key=lambda r: (r[0], r[1])):
        by_nk.setdefault((r[0], r[1]), []).append(r)
        by_n.setdefault(r[0], []).append(r)
    with open(out["Summary"], "w", newline="") as fs, open(out["RatioSummary"], "w", newline="") as fr:
        ws, wr = csv.writer(fs), csv.writer(fr)
        for (n, k), group in by_nk.items():
            names, t, wins, ratio, err = _stats(group)
            line, rline = [n, k], [n, k]
            for j, nm in enumerate(names):
                line += [round(t[:, j].mean(), 4), round(t[:, j].max(), 4), round(t[:, j].min(), 4),
                         int(wins[j])]
                rline += [nm, round(ratio[:, j].mean(), 4), round(ratio[:, j].max(), 4),
                          round(ratio[:, j].min(), 4)]
            ws.writerow(line + [len(group), err])
            wr.writerow(rline)
    with open(out["Overview"], "w", newline="") as fo, open(out["RatioOverview"], "w", newline="") as fq:
        wo, wq = csv.writer(fo), csv.writer(fq)
        for n, group in by_n.items():
            names, t, wins, ratio, err = _stats(group)
            line, rline = [n], [n]
            for j, nm in enumerate(names):
                line += [round(t[:, j].mean(), 4), round(t[:, j].max(), 4), round(t[:, j].min(), 4),
                         int(wins[j])]
                rline += [nm, round(ratio[:, j].mean(), 4), round(ratio[:, j].max(), 4),
                          round(ratio[:, j].min(), 4)]
            wo.writerow(line + [len(group), err])
            wq.writerow(rline)
    return out


# ------------------------------------------------------------------- CLI
def main(pkg, argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("crossover", help="numK crossover of bucket multiselect vs sort-and-choose")
    c.add_argument("--sizes", default="20,22,24", help="log2 vector sizes")
    c.add_argument("--dtype", default="f32", choices=["f32", "f64", "u32", "i32"])
    c.add_argument("--family", default="uniform")
    c.add_argument("--kfamily", default="uniform random", choices=K_FAMILIES)
    c.add_argument("--tests", type=int, default=3)
    c.add_argument("--csv", default=None)
    m = sub.add_parser("compare", help="time the algorithms on one problem")
    m.add_argument("--log2n", type=int, default=24)
    m.add_argument("--numk", type=int, default=1024)
    m.add_argument("--dtype", default="f32", choices=["f32", "f64", "u32", "i32"])
    m.add_argument("--family", default="uniform")
    m.add_argument("--kfamily", default="uniform random", choices=K_FAMILIES)
    m.add_argument("--tests", type=int, default=5)
    m.add_argument("--csv", default=None)
    s = sub.add_parser("summarize", help="Summary / Overview / Ratio files of result rows")
    s.add_argument("base")
    a = ap.parse_args(argv)
    dt = {"f32": np.float32, "f64": np.float64, "u32": np.uint32, "i32": np.int32}
    if a.cmd == "crossover":
        for lg in [int(x) for x in a.sizes.split(",")]:
            r = crossover(pkg, 1 << lg, dt[a.dtype], a.family, a.kfamily, a.tests,
                          csv_path=a.csv, log=print)
            print(f"n=2^{lg}: crossover numK = {r['crossover_num_k']} (bracket {r['bracket']})")
    elif a.cmd == "compare":
        r = compare(pkg, 1 << a.log2n, a.numk, dt[a.dtype], a.family, a.kfamily, a.tests,
                    algorithms=tuple(ALGORITHMS), csv_path=a.csv)
        print(r)
    else:
        for k, p in summarize(a.base).items():
            print(k, p)
    return 0


if __name__ == "__main__":
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from __graft_entry__ import load_package
    sys.exit(main(load_package()))

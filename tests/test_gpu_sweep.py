"""Seeded parity sweep over the compaction's size boundaries: rank counts
around 1024 children (the dense rounds' tables), 4096 (the LDS cursors of
4-byte keys) and beyond (cursor windows, row atomics), clustered and
uniform k, all key types and several distributions, bucket and radix
paths, each select against the oracle's sort-and-choose."""
import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu

FAMILIES = {np.float32: ["normal", "uniform", "cauchy", "plateau"],
            np.float64: ["normal", "uniform"], np.uint32: ["uniform"], np.int32: ["uniform"]}
RANKS = [900, 1024, 1100, 2100, 4000, 4200, 6000, 9000, 17000]


def _cases(count, seed=2024):
    rng = np.random.default_rng(seed)
    for case in range(count):
        dt = [np.float32, np.float64, np.uint32, np.int32][case % 4]
        fam = FAMILIES[dt][rng.integers(len(FAMILIES[dt]))]
        n = (1 << int(rng.integers(20, 23))) + int(rng.integers(0, 1000))
        nk = int(RANKS[rng.integers(len(RANKS))])
        kind = "radix" if rng.random() < 0.4 else "bucket"
        clustered = bool(rng.random() < 0.3)
        lo, hi = 0, 0
        if dt in (np.uint32, np.int32):
            lo, hi = (0, 2**31 - 1) if rng.random() < 0.5 else (0, 5000)
        c0 = int(rng.integers(1, max(2, n - nk)))
        k = (np.arange(c0, min(n, c0 + nk) + 1) if clustered
             else np.unique(rng.integers(1, n + 1, size=nk)))
        yield case, dt, fam, n, kind, lo, hi, k


@pytest.mark.parametrize("case", list(_cases(24)), ids=lambda c: f"case{c[0]}")
def test_boundary_sweep(pkg, O, case):
    i, dt, fam, n, kind, lo, hi, k = case
    print(f"case sweep {i}: {np.dtype(dt).name} {fam} n={n} {k.size} k {kind}", flush=True)
    d = pkg.generate(dt, fam, 1000 + i, n, lo=lo, hi=hi)
    x = d.to_host()
    ctx = pkg.Context()
    got = pkg.multiselect(d, k, ctx=ctx, kind=kind)
    assert np.array_equal(bits(got), bits(O.sort_choose(x, k)))
    ctx.close()
    d.free()

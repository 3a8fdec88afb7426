"""GPU parity at the benchmark configurations' full sizes, bit-exact against
the CPU oracle (its LSD radix sort-and-choose, fast enough at 2^28):

- n = 2^28 float32 N(0,1), 1024 uniform k (sorted, deduplicated), and a
  top-4096;
- n = 2^27 float64 N(0,1), same k distribution;
- n = 2^27 adversarial keys on the bucket and radix paths: uint32 in
  [0, 1023], float32 Cauchy(0,1) clipped to finite, float32 half 0.5 /
  half uniform; 256 k clustered in a width-1000 window at n/2, plus 1 and n.

Inputs come from the library's seeded device generator and are copied back
for the oracle.
"""
import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu


def device_vec(pkg, dtype, family, seed, n, lo=0, hi=0):
    d = pkg.generate(dtype, family, seed, n, lo=lo, hi=hi)
    return d, d.to_host()


def check(pkg, O, d, x, k, kind="bucket", label=""):
    print(f"case {label} kind={kind} n={x.size} k={len(k)}", flush=True)
    got = pkg.multiselect(d, k, kind=kind)
    want = O.sort_choose(x, k)
    bad = np.nonzero(bits(got) != bits(want))[0]
    assert bad.size == 0, f"{label} {kind}: {bad.size}/{len(k)} wrong, first k={np.asarray(k)[bad[:3]]}"


def test_f32_2p28_normal_1024k_and_topk(pkg, O):
    n = 1 << 28
    d, x = device_vec(pkg, np.float32, "normal", 1, n)
    k = pkg.workloads.k_uniform(101, n, 1024)
    check(pkg, O, d, x, k, label="f32 2^28 N(0,1)")
    print("case f32 2^28 top-4096", flush=True)
    assert np.array_equal(bits(pkg.topk(d, 4096)), bits(O.topk(x, 4096)))


def test_f64_2p27_normal_1024k(pkg, O):
    n = 1 << 27
    d, x = device_vec(pkg, np.float64, "normal", 2, n)
    check(pkg, O, d, x, pkg.workloads.k_uniform(101, n, 1024), label="f64 2^27 N(0,1)")


@pytest.mark.parametrize("case", ["u32_dup1024", "f32_cauchy", "f32_plateau"])
def test_adversarial_2p27_cluster_k(pkg, O, case):
    n = 1 << 27
    if case == "u32_dup1024":
        d, x = device_vec(pkg, np.uint32, "uniform", 3, n, 0, 1023)
    elif case == "f32_cauchy":
        d, x = device_vec(pkg, np.float32, "cauchy", 4, n)
        assert np.isfinite(x).all()
    else:
        d, x = device_vec(pkg, np.float32, "plateau", 5, n)
        assert 0.49 < (x == np.float32(0.5)).mean() < 0.51
    k = pkg.workloads.k_cluster(103, n)
    for kind in ("bucket", "radix"):
        check(pkg, O, d, x, k, kind=kind, label=case)

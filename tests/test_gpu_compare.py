"""GPU: the sort-and-choose entry point (the reference's baseline on the
device) against the oracle, the comparison driver's rows and crossover
search, and the single-vector tuning table (results never depend on it)."""
import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.uint32, np.int32])
def test_sort_choose_matches_oracle(pkg, O, dtype):
    W = pkg.workloads
    from ggks_multiselect_amd import _lib as L
    lib = L.lib()
    for n in (1, 1000, (1 << 20) + 3):
        x = {np.float32: W.normal_f32, np.float64: W.normal_f64}.get(dtype, None)
        x = x(5, n) if x else (W.uint_range(5, n, 0, 2**32 - 1) if dtype == np.uint32
                               else W.int_range(5, n, -2**31, 2**31 - 1))
        if dtype in (np.float32, np.float64) and n > 10:
            x[3], x[4] = -0.0, 0.0  # the radix order puts -0.0 first
        k = np.concatenate([W.k_uniform(6, n, 200, dedup=False), [1, n]]).astype(np.uint32)
        d = pkg.DeviceArray.from_host(x)
        out = np.empty(k.size, dtype)
        L.check(lib.ggks_sort_choose(None, pkg.api.dtype_code(dtype), d.ptr, n, k.ctypes.data,
                                     k.size, out.ctypes.data))
        assert np.array_equal(bits(out), bits(O.sort_choose(x, k))), n
        assert np.array_equal(bits(d.to_host()), bits(x))  # input intact


def test_compare_rows_and_crossover(pkg, tmp_path):
    C = pkg.compare
    path = str(tmp_path / "rows.csv")
    r = C.compare(pkg, 1 << 18, 64, np.float32, "normal", "cluster", tests=2,
                  algorithms=tuple(C.ALGORITHMS), csv_path=path)
    assert r["errors"] == 0 and sum(r["wins"].values()) == 2
    rows = C.read_rows([path])
    assert len(rows) == 2 and [a for a, _ in rows[0][4]] == list(C.ALGORITHMS)
    x = C.crossover(pkg, 1 << 16, tests=1, start=10, max_k=1 << 16, csv_path=path)
    assert x["probes"] and all(p["errors"] == 0 for p in x["probes"])
    out = C.summarize(str(tmp_path / "rows"))
    assert len(open(out["Summary"]).read().splitlines()) >= 2


def test_single_tuning_table(pkg, O):
    W = pkg.workloads
    n = (1 << 22) + 9
    x = W.normal_f32(8, n)
    k = W.k_uniform(9, n, 500)
    d = pkg.DeviceArray.from_host(x)
    want = O.sort_choose(x, k)
    try:
        for bl, ch in ((11, 64), (15, 512), (13, 200)):
            pkg.tuning_single_set(np.float32, 22, bl, ch)
            assert pkg.tuning_single_get(np.float32, 22) == (bl, ch)
            assert np.array_equal(bits(pkg.multiselect(d, k)), bits(want)), (bl, ch)
    finally:
        pkg.tuning_reset()
    assert pkg.tuning_single_get(np.float32, 22) == (12, 128)
    with pytest.raises(pkg.GGKSError):
        pkg.tuning_single_set(np.float32, 22, 16, 64)

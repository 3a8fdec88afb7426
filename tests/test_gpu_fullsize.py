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


def test_c4_batch_full_shape_every_segment(pkg, O):
    """The 4096-vector uneven batch at its full shape (lengths floor(2^U),
    U ~ U[10, 22], ~2.2e9 float32 keys, every band up to 2^21 present),
    graph replay, every segment checked against the oracle (its radix
    sort-and-choose or top-K, one C call per segment in a thread pool)."""
    from concurrent.futures import ThreadPoolExecutor
    spec = pkg.workloads.batch_spec(4096, 4096, 10, 22)
    table = pkg.segment_table(spec["lengths"], spec["kinds"], spec["num_k"])
    N = int(spec["lengths"].sum())
    bands = {int(np.log2(float(n))) for n in spec["lengths"]}
    assert 21 in bands and 10 in bands
    print(f"case c4 full shape: {len(table)} segments, {N} keys", flush=True)
    pkg.autotune.apply()
    d = pkg.DeviceArray(N, np.float32)
    for i, srow in enumerate(table):
        pkg.generate(np.float32, "normal" if spec["normal"][i] else "uniform", int(spec["seeds"][i]),
                     int(srow["length"]), out=d, offset=int(srow["offset"]))
    b = pkg.Batch(np.float32, table, spec["k"])
    got = b(d, graph=True)
    got2 = b(d, graph=True)
    assert np.array_equal(bits(got), bits(got2))
    x = d.to_host()
    k = spec["k"]

    def seg_ok(i):
        s = table[i]
        seg = x[int(s["offset"]):int(s["offset"] + s["length"])]
        if s["kind"] == 2:
            want = O.topk(seg, int(s["num_k"]))
        else:
            want = O.sort_choose(seg, k[int(s["k_offset"]):int(s["k_offset"] + s["num_k"])])
        g = got[int(s["out_offset"]):int(s["out_offset"] + s["num_k"])]
        return bool(np.array_equal(bits(g), bits(want)))

    with ThreadPoolExecutor(16) as ex:
        ok = list(ex.map(seg_ok, range(len(table))))
    bad = [i for i, v in enumerate(ok) if not v]
    assert not bad, f"{len(bad)} segments wrong, first {bad[:5]}"
    b.close()
    d.free()


def test_c5_pool_full_shape_trace(pkg):
    """A served trace over the serving workload's full pool (64 resident
    vectors, lengths log-uniform up to 2^26, f32 / u32 / f64), every answer
    checked against sorted host copies."""
    S = pkg.serving
    pool = S.make_pool(5000, 64)
    assert pool["lengths"].max() > 1 << 25
    bufs = S.upload_pool(pool)
    host = S.sorted_copies(pool, bufs)
    srv = pkg.Server(bufs, scratch_budget=32 << 30, max_inflight=4, max_batch=256)
    print("case c5 full pool, 8k req/s for 1 s", flush=True)
    r = S.run(srv, pool, host, 8000.0, 1.0, seed=5001)
    assert r["completed"] == r["requests"] > 5000 and r["bit_exact"], r
    srv.close()

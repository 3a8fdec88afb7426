"""CPU checks of the host-side logic around the kernels: segment tables,
batch and serving workloads, and the serving verifier (which must catch a
planted mistake)."""
import numpy as np

from conftest import bits


def test_segment_table_layout(pkg):
    t = pkg.segment_table([5, 7, 9], ["bucket", "topk", "radix"], [2, 3, 4])
    assert t["offset"].tolist() == [0, 5, 12]
    assert t["k_offset"].tolist() == [0, 2, 2]   # top-K takes no k values
    assert t["out_offset"].tolist() == [0, 2, 5]
    assert t["kind"].tolist() == [0, 2, 1]


def test_batch_spec_shape(pkg):
    s = pkg.workloads.batch_spec(4096, 4096, 10, 22)
    n = s["lengths"]
    assert n.min() >= 1024 and n.max() < 1 << 22
    assert 0.4 < (s["kinds"] == 2).mean() < 0.6
    ms = s["kinds"] != 2
    assert s["num_k"][ms].min() >= 1 and s["num_k"][ms].max() <= 64
    assert s["k"].size == int(s["num_k"][ms].sum())
    assert np.all(s["num_k"][~ms] <= np.minimum(n[~ms], 4096))
    # deterministic
    assert np.array_equal(pkg.workloads.batch_spec(4096, 50)["k"],
                          pkg.workloads.batch_spec(4096, 50)["k"])


def _fake_served(pkg, seed=3):
    """A small pool with host data, a trace and outputs filled by a direct
    per-request computation (a plain loop, independent of verify())."""
    S = pkg.serving
    W = pkg.workloads
    pool = S.make_pool(seed, 10, 6.0, 12.0)
    host = {}
    for j, dt in enumerate(S.DTYPES):
        host[dt] = np.zeros(pool["totals"][j], dt)
    for i in range(10):
        dt = S.DTYPES[pool["dt"][i]]
        o, n = int(pool["offsets"][i]), int(pool["lengths"][i])
        host[dt][o:o + n] = W.uint_range(i, n, 0, 50).astype(dt)  # duplicates on purpose
    srt = {dt: a.copy() for dt, a in host.items()}
    for i in range(10):
        dt = S.DTYPES[pool["dt"][i]]
        o, n = int(pool["offsets"][i]), int(pool["lengths"][i])
        srt[dt][o:o + n].sort()
    tr = S.make_trace(seed, pool, rate=2000.0, seconds=0.2)
    for i in range(tr["vec"].size):
        v = int(tr["vec"][i])
        dt = S.DTYPES[pool["dt"][v]]
        o, n = int(pool["offsets"][v]), int(pool["lengths"][v])
        desc = np.sort(host[dt][o:o + n])[::-1]
        nk = int(tr["num_k"][i])
        if tr["topk"][i]:
            vals = desc[:nk]
        else:
            ks = tr["kvals"][tr["koff"][i]:tr["koff"][i] + nk]
            vals = desc[ks.astype(np.int64) - 1]
        oo = int(tr["out_off"][i])
        tr["outs"][int(tr["dt"][i])][oo:oo + nk] = vals
    return pool, srt, tr


def test_serving_trace_and_verifier(pkg):
    S = pkg.serving
    pool, srt, tr = _fake_served(pkg)
    n = tr["arrival"].size
    assert 300 < n < 500 and np.all(np.diff(tr["arrival"]) >= 0) and tr["arrival"][-1] < 0.2
    r = tr["reqs"]
    assert np.all(r["num_k"] >= 1) and np.all(r["num_k"] <= np.maximum(256, 4096))
    assert np.all(r["num_k"][tr["topk"]] <= np.minimum(r["length"][tr["topk"]], 4096))
    assert S.verify(pool, srt, tr) == 0
    # a planted mistake: one output of one request off by one value
    j = int(tr["dt"][7])
    o = int(tr["out_off"][7])
    tr["outs"][j][o] = tr["outs"][j][o] + 1
    assert S.verify(pool, srt, tr) == 1


def test_pool_shape(pkg):
    S = pkg.serving
    pool = S.make_pool(5000, 64)
    n = pool["lengths"]
    assert n.min() >= 1 << 12 and n.max() < 1 << 26
    assert sum(pool["totals"]) == int(n.sum())
    share = np.bincount(pool["dt"], minlength=3) / 64.0
    assert share[0] > share[1] > share[2] > 0  # 6 : 3 : 1 in expectation
    assert bits(np.zeros(1, np.float32)).dtype == np.uint32

"""GPU tests of the serving loop (ggks_server_*): continuous batching of
requests over resident vectors, bit-exact against sorted host copies; each
answer is the same alone and in the mix; admission respects the scratch
budget."""
import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_pool(pkg):
    S = pkg.serving
    pool = S.make_pool(91, 24, 10.0, 19.0)
    bufs = S.upload_pool(pool)
    return pool, bufs, S.sorted_copies(pool, bufs, threads=4)


def test_trace_bit_exact_and_batched(pkg, small_pool):
    S = pkg.serving
    pool, bufs, host = small_pool
    srv = pkg.Server(bufs, scratch_budget=2 << 30, max_inflight=4, max_batch=64)
    res = S.run(srv, pool, host, rate=20000.0, seconds=0.5, seed=5)
    assert res["completed"] == res["requests"] > 1000
    assert res["bit_exact"], res
    st = srv.stats()
    assert st["pending"] == 0 and st["scratch_in_use"] == 0
    srv.close()
    # one slot and a burst far above what it serves one by one: requests that
    # arrive while a batch runs are grouped into the next
    srv = pkg.Server(bufs, scratch_budget=2 << 30, max_inflight=1, max_batch=64)
    res = S.run(srv, pool, host, rate=400000.0, seconds=0.02, seed=6)
    assert res["bit_exact"] and res["completed"] == res["requests"], res
    assert res["mean_batch"] > 4, res
    srv.close()


def test_alone_equals_in_the_mix(pkg, O, small_pool):
    S = pkg.serving
    pool, bufs, host = small_pool
    tr = S.make_trace(17, pool, rate=50000.0, seconds=0.02)
    srv = pkg.Server(bufs, scratch_budget=2 << 30, max_inflight=3, max_batch=32)
    lat = srv.run_trace(tr["reqs"], tr["arrival"])
    assert (lat >= 0).all()
    mixed = [o.copy() for o in tr["outs"]]
    assert S.verify(pool, host, tr) == 0
    # the same requests one at a time (nothing else in flight) and through the
    # single-vector entry points
    dtypes = S.DTYPES
    for i, r in enumerate(tr["reqs"]):
        j = int(tr["dt"][i])
        dt = dtypes[j]
        n, nk, o = int(r["length"]), int(r["num_k"]), int(tr["out_off"][i])
        out = np.zeros(nk, dt)
        view = bufs[dt].view(int(r["offset"]), n)
        if tr["topk"][i]:
            srv.submit(dt, int(r["offset"]), n, out, K=nk, kind="topk", tag=i)
            single = pkg.topk(view, nk)
        else:
            k = tr["kvals"][tr["koff"][i]:tr["koff"][i] + nk]
            srv.submit(dt, int(r["offset"]), n, out, k=k, tag=i)
            single = pkg.multiselect(view, k)
        srv.drain()
        done = srv.poll()
        assert [t for t, _ in done] == [i]
        assert np.array_equal(bits(out), bits(mixed[j][o:o + nk])), i
        assert np.array_equal(bits(out), bits(single)), i
    srv.close()


def test_admission_respects_budget(pkg, small_pool):
    S = pkg.serving
    pool, bufs, host = small_pool
    budget = 48 << 20
    srv = pkg.Server(bufs, scratch_budget=budget, max_inflight=4, max_batch=256)
    res = S.run(srv, pool, host, rate=30000.0, seconds=0.2, seed=8)
    assert res["bit_exact"] and res["completed"] == res["requests"]
    st = srv.stats()
    assert 0 < st["peak_scratch_bytes"] <= budget
    # one request needing more than the whole budget is refused up front
    dt = np.float32
    big = pkg.DeviceArray(1 << 24, dt)
    srv2 = pkg.Server({dt: big}, scratch_budget=16 << 20, max_inflight=1)
    with pytest.raises(pkg.GGKSError) as e:
        srv2.submit(dt, 0, 1 << 24, np.zeros(4, dt), k=[1, 2, 3, 4])
    assert e.value.code == pkg.api._lib.ENOMEM
    with pytest.raises(pkg.GGKSError):  # outside the resident buffer
        srv2.submit(dt, 10, 1 << 24, np.zeros(1, dt), k=[1])
    srv.close()
    srv2.close()

"""GPU parity of the second level: children of more than 2^16 keys are
re-bucketed by the whole GPU (k_count2 / k_compact2) instead of being
finished by one work-group.  It is on by default for radix plans, whose
equal-width digits leave heavy children on non-uniform data, and selected
with level2=1 for others.  Every output is checked against the oracle."""
import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _check(pkg, O, x, k, **kw):
    d = pkg.DeviceArray.from_host(x)
    got = pkg.multiselect(d, k, **kw)
    want = O.sort_choose(x, k)
    bad = np.nonzero(bits(got) != bits(want))[0]
    assert bad.size == 0, (kw, bad.size, bad[:4], got[bad[:4]], want[bad[:4]])
    d.free()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_radix_wide_range(pkg, O, dtype):
    """N(0,1) keys: the top digits hold ~1/300 of the vector each, far above
    2^16 keys, so most children go to the second level."""
    W = pkg.workloads
    n = (1 << 26) + 11 if dtype == np.float32 else (1 << 25) + 11
    x = W.normal_f32(21, n) if dtype == np.float32 else W.normal_f64(21, n)
    k = np.concatenate([W.k_uniform(22, n, 1024, dedup=False), W.k_cluster(23, n, 200, 600),
                        [1, 2, n - 1, n]]).astype(np.uint32)
    _say(f"radix wide range {np.dtype(dtype).name} n={n}")
    _check(pkg, O, x, k, kind="radix")
    _check(pkg, O, x, k, kind="radix", level2=-1)  # the same without it


def test_bucket_level2_spikes(pkg, O):
    """1 % spikes (below the heavy-value threshold) leave heavy buckets; with
    level2=1 they go to the second level."""
    W = pkg.workloads
    n = (1 << 24) + 5
    x = W.uniform_f32(31, n)
    z = W.splitmix64(32, 0, n)
    for j, v in enumerate((np.float32(0.25), np.float32(0.75), np.float32(0.999))):
        sel = (z % np.uint64(100)) == np.uint64(j)
        x[sel] = v
    k = np.concatenate([W.k_uniform(33, n, 500, dedup=False), [1, n, n // 4, n // 2]]).astype(np.uint32)
    _say(f"bucket level2 spikes n={n}")
    for l2 in (1, 0):
        _check(pkg, O, x, k, kind="bucket", level2=l2)
        _check(pkg, O, x, k, kind="radix", level2=l2)


def test_level2_topk_and_u32(pkg, O):
    W = pkg.workloads
    n = (1 << 24) + 3
    x = W.uniform_f32(41, n)
    z = W.splitmix64(42, 0, n)
    x[(z % np.uint64(50)) == 0] = np.float32(1.0)  # 2 % copies above every other key
    d = pkg.DeviceArray.from_host(x)
    for K in (1, 100, 5000):
        got = pkg.topk(d, K, level2=1)
        assert np.array_equal(bits(got), bits(O.topk(x, K))), K
    u = W.uint_range(43, n, 0, 2**32 - 1)
    u[(W.splitmix64(44, 0, n) % np.uint64(40)) == 0] = np.uint32(7)
    k = np.concatenate([W.k_uniform(45, n, 300, dedup=False), [1, n]]).astype(np.uint32)
    _check(pkg, O, u, k, kind="radix")
    _check(pkg, O, u, k, kind="bucket", level2=1)


def test_level2_batch(pkg, O):
    """Radix and bucket segments in one batch (the second level serves the
    radix ones by default, every segment with level2=1), graph replays."""
    W = pkg.workloads
    lens = [(1 << 23) + 1, 300000, (1 << 22) + 7, 5000, (1 << 21) + 3]
    kinds = ["radix", "bucket", "radix", "bucket", "topk"]
    xs = [W.normal_f32(50 + i, m) for i, m in enumerate(lens)]
    ks, nums = [], []
    for i, (m, kd) in enumerate(zip(lens, kinds)):
        if kd == "topk":
            nums.append(700)
        else:
            kk = np.concatenate([W.k_uniform(60 + i, m, 64, dedup=False), [1, m]]).astype(np.uint32)
            ks.append(kk)
            nums.append(kk.size)
    table = pkg.segment_table(lens, kinds, nums)
    d = pkg.DeviceArray.from_host(np.concatenate(xs))
    want = []
    j = 0
    for x, kd, nk in zip(xs, kinds, nums):
        if kd == "topk":
            want.append(O.topk(x, nk))
        else:
            want.append(O.sort_choose(x, ks[j]))
            j += 1
    want = np.concatenate(want)
    for l2 in (0, 1):
        b = pkg.Batch(np.float32, table, np.concatenate(ks), level2=l2)
        for rep in range(3):
            got = b(d, graph=True)
            assert np.array_equal(bits(got), bits(want)), (l2, rep)
        b.close()

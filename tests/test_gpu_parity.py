"""GPU parity: libggks.so (through its C ABI) against the CPU oracle, bit-exact.

Covers the four dtypes, the direct (small n) and engine paths, ragged and
misaligned lengths, unsorted and repeated k lists with k = 1 and k = n, +-0 and
+-inf, duplicate-heavy and heavy-tailed inputs, top-K, the radix path, the
stream-ordered entry point and the reference-named wrappers.
"""
import json
import os

import numpy as np
import pytest

from conftest import ROOT, bits

pytestmark = pytest.mark.gpu

DT = [np.float32, np.float64, np.uint32, np.int32]


def check(pkg, O, x, k, kind="bucket", **opt):
    d = pkg.DeviceArray.from_host(x)
    got = pkg.multiselect(d, k, kind=kind, **opt)
    want = O.sort_choose(x, k)
    bad = np.nonzero(bits(got) != bits(want))[0]
    assert bad.size == 0, (f"{kind} {x.dtype} n={x.size}: {bad.size}/{len(k)} wrong; first "
                           f"k={np.asarray(k)[bad[:3]]} got={got[bad[:3]]} want={want[bad[:3]]}")


def inputs(W, dtype, n, seed):
    if dtype == np.float32:
        return {"normal": W.normal_f32(seed, n), "uniform": W.uniform_f32(seed, n),
                "cauchy": W.cauchy_f32(seed, n), "plateau": W.plateau_f32(seed, n)}
    if dtype == np.float64:
        return {"normal": W.normal_f64(seed, n), "uniform": W.uniform_f64(seed, n)}
    if dtype == np.uint32:
        return {"full": W.uint_range(seed, n, 0, 2**32 - 1), "dup1024": W.uint_range(seed, n, 0, 1023),
                "mod4": (np.arange(n) % 4).astype(np.uint32)}
    return {"full": W.int_range(seed, n, -2**31, 2**31 - 1), "small": W.int_range(seed, n, -3, 3)}


def klist(W, n, seed, count=300):
    k = np.concatenate([W.k_uniform(seed, n, count, dedup=False), [1, n, n, 1, max(1, n // 2)]])
    return k.astype(np.uint32)


def test_config1_golden(pkg, O):
    with open(os.path.join(ROOT, "tests", "golden", "c1_uniform_f32_2p20.json")) as f:
        g = json.load(f)
    x = pkg.workloads.uniform_f32(0, g["n"])
    k = np.array(g["k"], np.uint32)
    want = np.array([int(h, 16) for h in g["expected_hex"]], np.uint32)
    d = pkg.DeviceArray.from_host(x)
    for kind in ("bucket", "radix"):
        assert np.array_equal(bits(pkg.multiselect(d, k, kind=kind)), want), kind
    out = np.zeros(k.size, np.float32)
    assert pkg.bucketMultiselectWrapper(d, g["n"], k, k.size, out, 256, 1024) == 1
    assert np.array_equal(bits(out), want)


@pytest.mark.parametrize("dtype", DT)
@pytest.mark.parametrize("log2n", [10, 14, 17, 20, 22])
def test_matrix(pkg, O, dtype, log2n):
    W = pkg.workloads
    n = (1 << log2n) + (log2n % 3)  # ragged lengths too
    for name, x in inputs(W, dtype, n, 100 + log2n).items():
        check(pkg, O, x, klist(W, n, log2n))


@pytest.mark.parametrize("dtype", DT)
def test_path_boundaries_and_tiny(pkg, O, dtype):
    W = pkg.workloads
    cap = 65536 // np.dtype(dtype).itemsize
    for n in (1, 2, 3, 7, 63, 64, 65, cap - 1, cap, cap + 1, 4 * cap + 3):
        x = inputs(W, dtype, n, n)[("normal" if dtype in (np.float32, np.float64) else "full")]
        k = np.arange(1, n + 1, dtype=np.uint32) if n <= 70 else klist(W, n, 7)
        check(pkg, O, x, k)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_signed_zero_and_infinities(pkg, O, dtype):
    n = 200_000
    W = pkg.workloads
    x = W.normal_f32(5, n).astype(dtype)
    x[::7] = -0.0
    x[3::7] = 0.0
    x[::1001] = np.inf
    x[5::1003] = -np.inf
    k = np.arange(1, n + 1, 97, dtype=np.uint32)
    check(pkg, O, x, np.concatenate([k, [1, n]]).astype(np.uint32))
    z = np.zeros(100_000, dtype)
    z[::2] = -0.0
    check(pkg, O, z, np.arange(1, 100_001, 999, dtype=np.uint32))


@pytest.mark.parametrize("dtype", DT)
def test_misaligned_views(pkg, O, dtype):
    W = pkg.workloads
    n = 300_001
    x = inputs(W, dtype, n, 77)[list(inputs(W, dtype, 8, 1))[0]]
    d = pkg.DeviceArray.from_host(x)
    k = klist(W, n - 10, 5)
    for off in (1, 2, 3):
        v = d.view(off, n - 10)
        got = pkg.multiselect(v, k)
        assert np.array_equal(bits(got), bits(O.sort_choose(x[off:off + n - 10], k)))


def test_all_equal_and_flat(pkg, O):
    for dtype in DT:
        x = np.full(1 << 20, 3, dtype)
        check(pkg, O, x, np.array([1, 2, 1 << 19, 1 << 20], np.uint32))


@pytest.mark.parametrize("dtype", DT)
def test_topk(pkg, O, dtype):
    W = pkg.workloads
    for n in (5, 1000, 20_000, 1 << 20):
        x = inputs(W, dtype, n, 3)[list(inputs(W, dtype, 8, 1))[0]]
        d = pkg.DeviceArray.from_host(x)
        for K in sorted({1, min(n, 10), min(n, 1000), min(n, 4096), n if n <= 1000 else 1}):
            got = pkg.topk(d, K)
            assert np.array_equal(bits(got), bits(O.topk(x, K))), (n, K)
    got = pkg.randomizedTopkSelectWrapper(d, d.n, 64)
    assert np.array_equal(bits(got), bits(O.topk(x, 64)))


@pytest.mark.parametrize("dtype", DT)
def test_radix_path(pkg, O, dtype):
    W = pkg.workloads
    for log2n in (12, 18, 21):
        n = (1 << log2n) + 5
        for name, x in inputs(W, dtype, n, log2n).items():
            check(pkg, O, x, klist(W, n, 3, count=64), kind="radix")
    x = W.normal_f32(8, 1 << 20)
    d = pkg.DeviceArray.from_host(x)
    assert bits(np.array([pkg.RadixSelectWrapper(d, d.n, 12345)], np.float32))[0] == \
        bits(O.sort_choose(x, [12345]))[0]


def test_async_entry_and_bad_k_status(pkg, O):
    from ggks_multiselect_amd import _lib as L
    W = pkg.workloads
    n = 1 << 21
    x = W.normal_f32(4, n)
    k = klist(W, n, 9)
    d = pkg.DeviceArray.from_host(x)
    dk = pkg.DeviceArray.from_host(k)
    dout = pkg.DeviceArray(k.size, np.float32)
    ctx = pkg.Context(0)
    for _ in range(3):  # repeated calls reuse the cached plan tables
        pkg.multiselect_async(d, dk, dout, ctx)
    pkg.sync()
    assert L.lib().ggks_ctx_check(ctx.handle) == L.OK
    assert np.array_equal(bits(dout.to_host()), bits(O.sort_choose(x, k)))
    badk = k.copy()
    badk[5] = n + 1
    dbad = pkg.DeviceArray.from_host(badk)
    pkg.multiselect_async(d, dbad, dout, ctx)
    pkg.sync()
    assert L.lib().ggks_ctx_check(ctx.handle) == L.EINVAL
    assert L.lib().ggks_ctx_check(ctx.handle) == L.OK  # cleared


def test_stage_events_and_marks(pkg, O):
    """Stage timing: all events give every stage and a whole-call time equal
    to their sum; a mask of the two events around the compaction pass gives
    that stage alone (the others read 0) and the same outputs.  The mask is
    latched when profiling is armed: changing it before the read changes
    nothing."""
    from ggks_multiselect_amd import _lib as L
    lib = L.lib()
    W = pkg.workloads
    n = 1 << 22
    x = W.normal_f32(6, n)
    k = klist(W, n, 6, count=64)
    d = pkg.DeviceArray.from_host(x)
    dk = pkg.DeviceArray.from_host(k)
    dout = pkg.DeviceArray(k.size, np.float32)
    ctx = pkg.Context(0)
    per = L.NSTAGES + 1
    ci = L.STAGES.index("compact")
    # a new context records every stage's events as markers: the stages sum
    # to the whole call
    L.check(lib.ggks_ctx_profile_arm(ctx.handle, 2))
    for _ in range(2):
        pkg.multiselect_async(d, dk, dout, ctx)
    ms = np.zeros(2 * per, np.float32)
    L.check(lib.ggks_ctx_profile_read(ctx.handle, ms.ctypes.data, 2))
    for row in ms.reshape(2, per):
        assert row[ci] > 0 and row[L.STAGES.index("count")] > 0
        assert abs(row[:L.NSTAGES].sum() - row[L.NSTAGES]) <= 1e-3 * row[L.NSTAGES] + 1e-3
    for mask in ((1 << per) - 1, (1 << ci) | (1 << (ci + 1))):
        L.check(lib.ggks_ctx_profile_marks(ctx.handle, mask))
        L.check(lib.ggks_ctx_profile_arm(ctx.handle, 2))
        for _ in range(2):
            pkg.multiselect_async(d, dk, dout, ctx)
        pkg.sync()
        L.check(lib.ggks_ctx_profile_marks(ctx.handle, (1 << per) - 1 - mask | 1))
        ms = np.zeros(2 * per, np.float32)
        L.check(lib.ggks_ctx_profile_read(ctx.handle, ms.ctypes.data, 2))
        ms = ms.reshape(2, per)
        assert np.array_equal(bits(dout.to_host()), bits(O.sort_choose(x, k)))
        for row in ms:
            if mask == (1 << per) - 1:
                assert row[ci] > 0 and row[L.STAGES.index("count")] > 0
                assert abs(row[:L.NSTAGES].sum() - row[L.NSTAGES]) <= 1e-3 * row[L.NSTAGES] + 1e-3
            else:
                assert row[ci] > 0 and row[L.NSTAGES] == row[ci]
                assert not np.any(np.delete(row[:L.NSTAGES], ci))
    L.check(lib.ggks_ctx_profile_arm(ctx.handle, 0))
    L.check(lib.ggks_ctx_profile_marks(ctx.handle, (1 << per) - 1))


def test_kernel_events(pkg, O):
    """Kernel-attached events: with GGKS_PROFILE_KERNEL_EVENTS the count and
    compaction kernels' launches carry the two events of their stage.  Each
    stage then reads a positive time no longer than the whole call measured
    by markers, the other stages read 0, and the outputs are unchanged."""
    from ggks_multiselect_amd import _lib as L
    lib = L.lib()
    W = pkg.workloads
    n = 1 << 22
    x = W.normal_f32(7, n)
    k = klist(W, n, 7, count=64)
    d = pkg.DeviceArray.from_host(x)
    dk = pkg.DeviceArray.from_host(k)
    dout = pkg.DeviceArray(k.size, np.float32)
    ctx = pkg.Context(0)
    per = L.NSTAGES + 1
    want = bits(O.sort_choose(x, k))
    # the whole call by markers, for the bound
    L.check(lib.ggks_ctx_profile_marks(ctx.handle, (1 << per) - 1))
    L.check(lib.ggks_ctx_profile_arm(ctx.handle, 3))
    for _ in range(3):
        pkg.multiselect_async(d, dk, dout, ctx)
    ms = np.zeros(3 * per, np.float32)
    L.check(lib.ggks_ctx_profile_read(ctx.handle, ms.ctypes.data, 3))
    whole = float(ms.reshape(3, per)[:, L.NSTAGES].max())
    for name in ("count", "compact"):
        si = L.STAGES.index(name)
        L.check(lib.ggks_ctx_profile_marks(ctx.handle,
                                           (1 << si) | (1 << (si + 1)) | L.PROFILE_KERNEL_EVENTS))
        L.check(lib.ggks_ctx_profile_arm(ctx.handle, 3))
        for _ in range(3):
            pkg.multiselect_async(d, dk, dout, ctx)
        ms = np.zeros(3 * per, np.float32)
        L.check(lib.ggks_ctx_profile_read(ctx.handle, ms.ctypes.data, 3))
        assert np.array_equal(bits(dout.to_host()), want)
        for row in ms.reshape(3, per):
            assert 0 < row[si] <= whole, (name, row[si], whole)
            assert not np.any(np.delete(row[:L.NSTAGES], si))
    L.check(lib.ggks_ctx_profile_arm(ctx.handle, 0))
    L.check(lib.ggks_ctx_profile_marks(ctx.handle, (1 << per) - 1))


def test_options_do_not_change_results(pkg, O):
    W = pkg.workloads
    n = (1 << 22) + 1
    x = W.normal_f32(12, n)
    k = klist(W, n, 12, count=1000)
    for opt in ({"buckets_log2": 8}, {"buckets_log2": 15}, {"blocks": 1}, {"blocks": 37},
                {"seed": 99}, {"seed": 12345678901}):
        check(pkg, O, x, k, **opt)
    for rb in (4, 8, 11, 14):
        check(pkg, O, x, k, kind="radix", radix_bits=rb)


def test_device_copy_plumbing(pkg):
    from ggks_multiselect_amd import _lib as L
    W = pkg.workloads
    x = W.uint_range(21, (1 << 20) + 7, 0, 2**32 - 1)
    d = pkg.DeviceArray.from_host(x)
    for nbytes in (16, 4096, (1 << 20) * 4, 4 * x.size, 4 * x.size - 12):  # kernel and runtime paths
        y = pkg.DeviceArray(x.size, np.uint32)
        L.check(L.lib().ggks_dev_fill(y.ptr, 0, 4 * x.size))
        L.check(L.lib().ggks_d2d(y.ptr, d.ptr, nbytes))
        h = y.to_host()
        m = nbytes // 4
        assert np.array_equal(h[:m], x[:m]) and not np.any(h[m:])


def test_cpp_compat_harness():
    """The reference-named C++ wrappers (include/ggks_compat.hpp) through a
    g++-built harness that checks them against std::sort on host copies."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "cpp", "compat_harness")
    assert os.path.exists(exe), "built by ggks-multiselect_amd/build.py"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr

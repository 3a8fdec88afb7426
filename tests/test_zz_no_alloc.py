"""GPU: the no-allocation entry point (ggks_bucket_multiselect_noalloc)
through the C ABI against the stream-ordered context entry point
(ggks_bucket_multiselect_async) on the same device data, bit-exact, on
calls in a row whose inputs shrink (engine path, then a smaller engine
select, then the direct path), for each key type; and its size limit."""
import ctypes

import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu


def _run_pair(pkg, L, x, k, ctx):
    lib = L.lib()
    code = pkg.api.dtype_code(x.dtype)
    d = pkg.DeviceArray.from_host(x)
    dk = pkg.DeviceArray.from_host(k)
    out_ref = pkg.DeviceArray(k.size, x.dtype)
    out_na = pkg.DeviceArray(k.size, x.dtype)
    L.check(lib.ggks_bucket_multiselect_async(ctx.handle, code, d.ptr, x.size, dk.ptr, k.size,
                                              out_ref.ptr, None))
    L.check(lib.ggks_bucket_multiselect_noalloc(code, d.ptr, x.size, dk.ptr, k.size, out_na.ptr,
                                                None, None))
    L.check(lib.ggks_sync())
    L.check(lib.ggks_ctx_check(ctx.handle))
    return out_ref.to_host(), out_na.to_host()


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.uint32, np.int32])
def test_noalloc_matches_async(pkg, O, dtype):
    from ggks_multiselect_amd import _lib as L
    W = pkg.workloads
    ctx = pkg.Context(0)
    big = (1 << 22) + 5 if np.dtype(dtype).itemsize == 4 else (1 << 21) + 5
    for n, seed in ((big, 1), (20000, 2), (3001, 3)):  # each call smaller than the last
        print(f"noalloc {np.dtype(dtype).name} n={n}", flush=True)
        if dtype == np.float32:
            x = W.normal_f32(seed, n)
        elif dtype == np.float64:
            x = W.normal_f64(seed, n)
        elif dtype == np.uint32:
            x = W.uint_range(seed, n, 0, 1023 if seed == 2 else 2**32 - 1)
        else:
            x = W.int_range(seed, n, -2**31, 2**31 - 1)
        k = np.concatenate([W.k_uniform(seed + 10, n, 300, dedup=False),
                            [1, n, n // 2]]).astype(np.uint32)
        ref, na = _run_pair(pkg, L, x, k, ctx)
        assert np.array_equal(bits(na), bits(ref)), (n, np.nonzero(bits(na) != bits(ref))[0][:5])
        assert np.array_equal(bits(na), bits(O.sort_choose(x, k))), n
    ctx.close()


def test_noalloc_limit(pkg):
    from ggks_multiselect_amd import _lib as L
    lib = L.lib()
    d = pkg.DeviceArray(1 << 25, np.float32)
    dk = pkg.DeviceArray.from_host(np.array([1, 2], np.uint32))
    out = pkg.DeviceArray(2, np.float32)
    rc = lib.ggks_bucket_multiselect_noalloc(L.F32, d.ptr, 1 << 25, dk.ptr, 2, out.ptr, None, None)
    assert rc == L.ENOMEM and b"no-allocation" in lib.ggks_last_error()
    # nothing to do, bad arguments: refused before any device work
    assert lib.ggks_bucket_multiselect_noalloc(L.F32, d.ptr, 10, None, 0, None, None, None) == L.OK
    assert lib.ggks_bucket_multiselect_noalloc(L.F32, None, 10, dk.ptr, 2, out.ptr, None,
                                               None) == L.EINVAL

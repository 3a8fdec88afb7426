"""CPU-side checks of the C ABI: libggks.so loads, exports every function
include/ggks.h declares, and rejects bad arguments before touching a device."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def header_functions():
    with open(os.path.join(ROOT, "include", "ggks.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ggks_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def L(pkg):
    from ggks_multiselect_amd import _lib
    return _lib


def test_library_exports_header(L):
    lib = L.lib()
    names = header_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(L.EXPORTS) == names


def test_version_and_scratch_model(L):
    lib = L.lib()
    assert b"gfx950" in lib.ggks_version()
    s1 = lib.ggks_scratch_bytes(L.KIND_BUCKET, L.F32, 1 << 20, 101)
    s2 = lib.ggks_scratch_bytes(L.KIND_BUCKET, L.F32, 1 << 28, 1024)
    s3 = lib.ggks_scratch_bytes(L.KIND_BUCKET, L.F64, 1 << 28, 1024)
    # worst case keeps a reduced copy of the vector: at least n * sizeof(T)
    assert s1 >= (1 << 20) * 4 and s2 >= (1 << 28) * 4 and s3 >= (1 << 28) * 8
    assert s2 < (1 << 28) * 4 * 1.1  # plus small tables (rows, LUT, children)
    assert lib.ggks_scratch_bytes(L.KIND_BUCKET, 9, 100, 1) == 0


def test_bad_arguments_rejected_without_device(L):
    lib = L.lib()
    k = np.array([1, 2], np.uint32)
    out = np.zeros(2, np.float32)
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: validation fails first
    cases = [
        (L.F32, fake, 0, k, 2),          # n < 1
        (L.F32, fake, 1 << 40, k, 2),    # n >= 2^32
        (7, fake, 10, k, 2),             # bad dtype
        (L.F32, None, 10, k, 2),         # null vector
        (L.F32, fake, 10, np.array([0, 1], np.uint32), 2),   # k = 0
        (L.F32, fake, 10, np.array([11, 1], np.uint32), 2),  # k > n
    ]
    for dt, vec, n, kk, nk in cases:
        rc = lib.ggks_bucket_multiselect(None, dt, vec, n, kk.ctypes.data, nk,
                                         out.ctypes.data, None)
        assert rc == L.EINVAL, (dt, n, kk)
        assert lib.ggks_last_error()
        rc = lib.ggks_radix_multiselect(None, dt, vec, n, kk.ctypes.data, nk,
                                        out.ctypes.data, None)
        assert rc == L.EINVAL
    assert lib.ggks_topk(None, L.F32, fake, 10, 0, out.ctypes.data, None) == L.EINVAL
    assert lib.ggks_topk(None, L.F32, fake, 10, 11, out.ctypes.data, None) == L.EINVAL
    # nothing to do: no device needed either
    assert lib.ggks_bucket_multiselect(None, L.F32, fake, 10, None, 0, None, None) == L.OK
    assert np.all(out == 0)  # outputs untouched on error


def test_python_wrappers_validate(pkg):
    with pytest.raises(pkg.GGKSError):
        pkg.api._k_array([0, 3])
    with pytest.raises(TypeError):
        pkg.api.dtype_code(np.int16)


def test_workload_generators_are_pinned(pkg):
    W = pkg.workloads
    # SplitMix64 reference outputs (seed 0, first three), from its definition
    z = W.splitmix64(0, 0, 3)
    assert [hex(int(v)) for v in z] == ["0xe220a8397b1dcdaf", "0x6e789e6aa1b965f4",
                                        "0x6c45d188009454f"]
    x = W.normal_f32(1, 1 << 16)
    assert abs(float(x.mean())) < 0.02 and abs(float(x.std()) - 1) < 0.02
    assert not np.any(np.signbit(x) & (x == 0))
    p = W.plateau_f32(5, 1000)
    assert int(np.sum(p == np.float32(0.5))) >= 500
    kc = W.k_cluster(103, 1 << 27)
    assert kc.size == 258 and kc[-2] == 1 and kc[-1] == 1 << 27
    assert kc[:-2].min() >= (1 << 26) - 500 and kc[:-2].max() <= (1 << 26) + 499
    assert np.array_equal(W.k_percentiles(101), np.arange(1, 102)[[round(p) for p in
                                                                   np.arange(101) * 1.0]])

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


def test_batch_validation_without_device(pkg, L):
    lib = L.lib()
    table = pkg.segment_table([100, 50, 10], ["bucket", "topk", "radix"], [3, 50, 2])
    k = np.array([1, 100, 7, 10, 1], np.uint32)
    assert pkg.Batch.scratch_bytes(np.float32, table, k) > 100 * 4
    import ctypes
    for bad_k, bad_table in (
            (np.array([1, 101, 7, 10, 1], np.uint32), table),       # k > length
            (np.array([0, 100, 7, 10, 1], np.uint32), table),       # k = 0
            (k[:4], table)):                                         # k list too short
        h = ctypes.c_void_p()
        rc = lib.ggks_batch_create(ctypes.byref(h), 0, L.F32, bad_table.ctypes.data, bad_table.size,
                                   bad_k.ctypes.data, bad_k.size, None, None, 0, None)
        assert rc == L.EINVAL and lib.ggks_last_error()
    t2 = table.copy()
    t2["num_k"][1] = 51  # K > length
    t3 = table.copy()
    t3["kind"][0] = 9
    t4 = table.copy()
    t4["length"][2] = 0
    for t in (t2, t3, t4):
        assert pkg.Batch.scratch_bytes(np.float32, t, k) == 0
    with pytest.raises(pkg.GGKSError):
        pkg.tuning_set(np.float32, 40, 12, 16, False)
    with pytest.raises(pkg.GGKSError):
        pkg.tuning_set(np.float32, 20, 16, 16, False)
    pkg.tuning_set(np.float32, 20, 12, 16, False)
    assert pkg.tuning_get(np.float32, 20) == (12, 16, False)
    pkg.tuning_reset()
    assert pkg.tuning_get(np.float32, 12)[2] is True   # 2^12..2^13-1 f32 fit in LDS: direct
    assert pkg.tuning_get(np.float32, 14)[2] is False  # up to 2^15-1 does not

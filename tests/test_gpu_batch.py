"""GPU parity of the batched segment-table path (libggks.so, ggks_batch_*)
against the CPU oracle, bit-exact per segment: uneven lengths across the
direct / engine boundary, multiselect and top-K segments mixed, every dtype,
graph replays, tuning changes and agreement with the single-vector calls."""
import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu


def host_values(W, dtype, spec):
    parts = []
    for n, sd, nm in zip(spec["lengths"], spec["seeds"], spec["normal"]):
        n, sd = int(n), int(sd)
        if dtype == np.float32:
            parts.append((W.normal_f32 if nm else W.uniform_f32)(sd, n))
        elif dtype == np.float64:
            parts.append((W.normal_f64 if nm else W.uniform_f64)(sd, n))
        elif dtype == np.uint32:  # duplicate-heavy or full range by the coin
            parts.append(W.uint_range(sd, n, 0, 1023 if nm else 2**32 - 1))
        else:
            parts.append(W.int_range(sd, n, -5 if nm else -2**31, 5 if nm else 2**31 - 1))
    return np.concatenate(parts)


def expected(O, x, table, k):
    want = []
    for s in table:
        seg = x[int(s["offset"]):int(s["offset"] + s["length"])]
        if s["kind"] == 2:
            want.append(O.topk(seg, int(s["num_k"])))
        else:
            want.append(O.sort_choose(seg, k[int(s["k_offset"]):int(s["k_offset"] + s["num_k"])]))
    return want


def check_outputs(got, want, table):
    for i, (s, w) in enumerate(zip(table, want)):
        g = got[int(s["out_offset"]):int(s["out_offset"] + s["num_k"])]
        bad = np.nonzero(bits(g) != bits(w))[0]
        assert bad.size == 0, (f"segment {i} len={int(s['length'])} kind={int(s['kind'])}: "
                               f"{bad.size}/{len(w)} wrong, first at {bad[:3]}")


def make(pkg, dtype, seed, nseg, lo, hi):
    W = pkg.workloads
    spec = W.batch_spec(seed, nseg, lo, hi)
    table = pkg.segment_table(spec["lengths"], spec["kinds"], spec["num_k"])
    x = host_values(W, dtype, spec)
    return spec, table, x


@pytest.mark.parametrize("dtype,nseg,lo,hi", [(np.float32, 300, 6, 20), (np.float64, 120, 6, 19),
                                              (np.uint32, 120, 6, 19), (np.int32, 120, 6, 19)])
def test_batch_matches_oracle(pkg, O, dtype, nseg, lo, hi):
    spec, table, x = make(pkg, dtype, 700 + nseg, nseg, lo, hi)
    d = pkg.DeviceArray.from_host(x)
    b = pkg.Batch(dtype, table, spec["k"])
    st = b.stats()
    assert st["engine_segments"] > 0 and st["direct_segments"] > 0
    assert st["queries"] == int(spec["num_k"].sum())
    check_outputs(b(d), expected(O, x, table, spec["k"]), table)


def test_batch_many_children(pkg, O):
    # more than 256 x 1024 queries: every finisher work-group screens its
    # share of the child table in several windows
    W = pkg.workloads
    spec = W.batch_spec(4242, 1500, 10, 14, max_k=64, max_topk=4096, topk_frac=0.9)
    table = pkg.segment_table(spec["lengths"], spec["kinds"], spec["num_k"])
    assert int(spec["num_k"].sum()) > 256 * 1024
    x = host_values(W, np.float32, spec)
    d = pkg.DeviceArray.from_host(x)
    b = pkg.Batch(np.float32, table, spec["k"])
    check_outputs(b(d), expected(O, x, table, spec["k"]), table)


def test_batch_large_child_small_buckets(pkg, O):
    # segments whose mean bucket is small (the register-light finisher is
    # chosen) but which hold a cluster of three adjacent values ~4000 times
    # each: the bucket holding it is a child of ~12000 keys, past that
    # finisher's registers, with up to four ranks inside it
    W = pkg.workloads
    n, nseg = 1 << 16, 4
    xs, ks = [], []
    for i in range(nseg):
        x = W.uniform_f32(900 + i, n)
        v = np.float32(0.125 * (i + 1))
        cl = np.array([v, np.nextafter(v, np.float32(2)),
                       np.nextafter(np.nextafter(v, np.float32(2)), np.float32(2))], np.float32)
        pos = np.sort(W.k_uniform(950 + i, n, 12000, dedup=True).astype(np.int64) - 1)
        x[pos] = cl[np.arange(pos.size) % 3]
        s = np.sort(x)
        first, last = np.searchsorted(s, cl[0]), np.searchsorted(s, cl[2], side="right")
        k = n - np.array([first + 5, (first + last) // 2, last - 5, first + (last - first) // 3][:1 + i])
        ks.append(np.concatenate([k, [1 + i * 1000]]).astype(np.uint32))
        xs.append(x)
    table = pkg.segment_table([n] * nseg, [0] * nseg, [k.size for k in ks])
    x, k = np.concatenate(xs), np.concatenate(ks)
    b = pkg.Batch(np.float32, table, k)
    assert b.stats()["engine_segments"] == nseg
    d = pkg.DeviceArray.from_host(x)
    check_outputs(b(d), expected(O, x, table, k), table)
    check_outputs(b(d, graph=True), expected(O, x, table, k), table)


def test_batch_graph_replays(pkg, O):
    spec, table, x = make(pkg, np.float32, 11, 200, 8, 21)
    d = pkg.DeviceArray.from_host(x)
    b = pkg.Batch(np.float32, table, spec["k"])
    eager = b(d)
    check_outputs(eager, expected(O, x, table, spec["k"]), table)
    out = pkg.DeviceArray(b.num_outputs, np.float32)
    for rep in range(4):
        pkg.api.check(pkg.api.lib().ggks_dev_fill(out.ptr, 0xFF, out.nbytes))
        b.run(d, out, graph=True)
        b.sync()
        assert b.stats()["graph"] == 1
        assert np.array_equal(bits(out.to_host()), bits(eager)), rep
        # eager work of the program between replays
        assert np.array_equal(bits(b(d)), bits(eager))


def test_batch_equals_single_calls(pkg, O):
    spec, table, x = make(pkg, np.float32, 21, 60, 10, 21)
    d = pkg.DeviceArray.from_host(x)
    got = pkg.Batch(np.float32, table, spec["k"])(d)
    k = spec["k"]
    for s in table:
        v = d.view(int(s["offset"]), int(s["length"]))
        if s["kind"] == 2:
            one = pkg.topk(v, int(s["num_k"]))
        else:
            one = pkg.multiselect(v, k[int(s["k_offset"]):int(s["k_offset"] + s["num_k"])])
        g = got[int(s["out_offset"]):int(s["out_offset"] + s["num_k"])]
        assert np.array_equal(bits(g), bits(one))


def test_batch_tuning_does_not_change_results(pkg, O):
    spec, table, x = make(pkg, np.float32, 31, 150, 10, 20)
    d = pkg.DeviceArray.from_host(x)
    want = expected(O, x, table, spec["k"])
    try:
        for bl, cl, direct in ((8, 12, False), (15, 18, False), (11, 14, True)):
            for band in range(10, 21):
                pkg.tuning_set(np.float32, band, bl, cl, direct)
            check_outputs(pkg.Batch(np.float32, table, spec["k"])(d), want, table)
    finally:
        pkg.tuning_reset()


def test_batch_radix_segments_and_ragged_offsets(pkg, O):
    W = pkg.workloads
    lens = np.array([1, 5, 64, 65, 5000, 70_001, 300_000, 1 << 20], np.uint64)
    gaps = np.array([3, 0, 7, 1, 0, 13, 2, 5], np.uint64)  # unaligned, gapped layout
    offs = np.cumsum(gaps + np.concatenate([[0], lens[:-1]])).astype(np.uint64)
    x = W.normal_f32(5, int(offs[-1] + lens[-1]) + 3)
    kinds = ["radix", "bucket", "topk", "radix", "topk", "radix", "bucket", "radix"]
    nk = np.array([1, 5, 64, 30, 4096, 64, 17, 64], np.uint32)
    ks, koffs, pos = [], [], 0
    for n, kd, c in zip(lens, kinds, nk):
        koffs.append(pos)
        if kd != "topk":
            ks.append((W.splitmix64(int(n), 0, int(c)) % n + 1).astype(np.uint32))
            pos += int(c)
    k = np.concatenate(ks)
    table = pkg.segment_table(lens, kinds, nk, offsets=offs, k_offsets=np.array(koffs, np.uint64))
    d = pkg.DeviceArray.from_host(x)
    check_outputs(pkg.Batch(np.float32, table, k)(d), expected(O, x, table, k), table)


def test_device_generator_matches_host(pkg):
    W = pkg.workloads
    n = 100_003
    for dt, fam, lo, hi, host in (
            (np.float32, "uniform", 0, 0, lambda: W.uniform_f32(9, n + 7)[7:]),
            (np.float64, "uniform", 0, 0, lambda: W.uniform_f64(9, n + 7)[7:]),
            (np.uint32, "uniform", 0, 1023, lambda: W.uint_range(9, n + 7, 0, 1023)[7:]),
            (np.int32, "uniform", -2**31, 2**31 - 1,
             lambda: W.int_range(9, n + 7, -2**31, 2**31 - 1)[7:])):
        got = pkg.generate(dt, fam, 9, n, start=7, lo=lo, hi=hi).to_host()
        assert np.array_equal(bits(got) if dt != np.int32 else got, bits(host()) if dt != np.int32
                              else host()), dt
    g = pkg.generate(np.float32, "normal", 1, n).to_host()
    h = W.normal_f32(1, n)
    ulp = np.abs(g.view(np.int32).astype(np.int64) - h.view(np.int32).astype(np.int64))
    assert ulp.max() <= 1 and (ulp == 0).mean() > 0.99
    assert not np.any((g == 0) & np.signbit(g))

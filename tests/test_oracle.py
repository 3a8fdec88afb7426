"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each oracle function is checked against something other than itself:
  key conversion   -- exhaustively: a strictly increasing bijection along the
                      value order of every float32 bit pattern; f64 specials and
                      random pairs against IEEE comparison
  sort_choose      -- brute force on tiny inputs against an order written from
                      its definition; closed-form known answers (permutation,
                      all-equal, ones/twos, i % 4, duplicate-heavy histogram,
                      bucket-killer powers of two); the committed config-1 golden
  topk             -- brute force; the reverse of sort_choose
  radix_select     -- sort_choose, and its digits are the answer's key nibbles
  bucket_multiselect (restatement) -- sort_choose on many inputs, plus the
                      invariants sum(counts) = n and monotone bucketing
  randomized_topk  -- topk
  nth / parallel   -- sort_choose
  oracle/_ref      -- sort_choose (when the reference build exists)
and a planted mistake shows the comparisons can fail.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import ROOT, bits, total_order_sort

DT = [np.float32, np.float64, np.uint32, np.int32]


def tiny_vectors(rng, dtype, n):
    if dtype in (np.float32, np.float64):
        pool = np.array([-np.inf, -3.5, -1.0, -0.0, 0.0, 1e-30, 0.5, 2.0, 7.25, np.inf], dtype)
        return rng.choice(pool, n)
    if dtype == np.uint32:
        return rng.choice(np.array([0, 1, 2, 7, 2**31, 2**32 - 1], np.uint32), n)
    return rng.choice(np.array([-2**31, -5, -1, 0, 1, 3, 2**31 - 1], np.int32), n)


# ---------------------------------------------------------------- keys
def test_keyconv_f32_exhaustive_monotone_bijection(O):
    """Every one of the 2^32 float32 patterns round-trips; along the value
    order of all non-NaN floats the keys strictly increase: non-negative
    floats ascend with their bits, negative floats descend with theirs, and
    -0.0 sits directly below +0.0."""
    chunk = 1 << 26
    POS_END, NEG0, NEG_END = 0x7F800001, 0x80000000, 0xFF800001  # +inf + 1, -0.0, -inf + 1
    last_pos = last_neg = None
    for s in range(0, 1 << 32, chunk):
        b = np.arange(s, s + chunk, dtype=np.uint64).astype(np.uint32)
        k = O.keys(b.view(np.float32))
        assert np.array_equal(O.unkeys(k, np.float32).view(np.uint32), b)
        if s + chunk <= POS_END or s < POS_END:
            kp = k[: max(0, min(chunk, POS_END - s))]
            assert np.all(kp[1:] > kp[:-1])
            if kp.size:
                assert last_pos is None or kp[0] > last_pos
                last_pos = kp[-1]
        if s >= NEG0 and s < NEG_END:
            kn = k[: min(chunk, NEG_END - s)]
            assert np.all(kn[1:] < kn[:-1])  # more negative as the bits grow
            assert last_neg is None or kn[0] < last_neg
            last_neg = kn[-1]
    kz = O.keys(np.array([0x80000000, 0], np.uint32).view(np.float32))
    assert int(kz[0]) + 1 == int(kz[1])  # -0.0 immediately below +0.0
    kinf = O.keys(np.array([-np.inf, np.inf], np.float32))
    assert int(kinf[0]) == int(last_neg) and int(kinf[1]) == int(last_pos)


def test_keyconv_f64_specials_and_pairs(O):
    sp = np.array([-np.inf, -np.finfo(np.float64).max, -1.0, -np.finfo(np.float64).tiny,
                   -5e-324, -0.0, 0.0, 5e-324, np.finfo(np.float64).tiny, 1.0,
                   np.finfo(np.float64).max, np.inf])
    k = O.keys(sp)
    assert np.all(np.diff(k.astype(object)) > 0)
    rng = np.random.default_rng(7)
    a = rng.integers(0, 2**64 - 1, 2_000_000, dtype=np.uint64, endpoint=True).view(np.float64)
    b = rng.integers(0, 2**64 - 1, 2_000_000, dtype=np.uint64, endpoint=True).view(np.float64)
    ok = ~(np.isnan(a) | np.isnan(b) | ((a == 0) & (b == 0)))
    a, b = a[ok], b[ok]
    ka, kb = O.keys(a), O.keys(b)
    assert np.array_equal(ka < kb, a < b)
    assert np.array_equal(ka == kb, bits(a) == bits(b))
    assert np.array_equal(bits(O.unkeys(ka, np.float64)), bits(a))


def test_keyconv_ints(O):
    i = np.array([-2**31, -1, 0, 1, 2**31 - 1], np.int32)
    assert np.all(np.diff(O.keys(i).astype(np.int64)) > 0)
    rng = np.random.default_rng(3)
    a = rng.integers(-2**31, 2**31, 1_000_000, dtype=np.int64).astype(np.int32)
    b = rng.integers(-2**31, 2**31, 1_000_000, dtype=np.int64).astype(np.int32)
    assert np.array_equal(O.keys(a) < O.keys(b), a < b)
    u = rng.integers(0, 2**32, 1000, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(O.keys(u), u)


# --------------------------------------------------------- sort & choose
@pytest.mark.parametrize("dtype", DT)
def test_sort_choose_bruteforce_tiny(O, dtype):
    rng = np.random.default_rng(11)
    for n in list(range(1, 20)) + [31, 64]:
        for _ in range(4):
            x = tiny_vectors(rng, dtype, n).astype(dtype)
            s = total_order_sort(x)
            k = np.arange(1, n + 1, dtype=np.uint32)
            rng.shuffle(k)
            got = O.sort_choose(x, k)
            assert np.array_equal(bits(got), bits(s[n - k.astype(np.int64)]))
            assert np.array_equal(bits(O.sort(x)), bits(s))


@pytest.mark.parametrize("dtype", DT)
def test_sort_choose_permutation_kat(O, dtype, pkg):
    n = 1 << 16
    x = pkg.workloads.permutation_u32(5, n).astype(dtype)
    k = np.random.default_rng(1).integers(1, n + 1, 2000).astype(np.uint32)
    assert np.array_equal(O.sort_choose(x, k), (n - k.astype(np.int64) + 1).astype(dtype))


def test_sort_choose_closed_forms(O, pkg):
    W = pkg.workloads
    n = 100_000
    # all equal
    x = np.full(n, 1.0, np.float32)
    assert np.all(O.sort_choose(x, [1, 2, n // 2, n]) == 1.0)
    # ones and twos, 50/50 (generateProblems.cu:154-177 shape)
    x = np.where(np.arange(n) % 2 == 0, 2.0, 1.0).astype(np.float32)
    np.random.default_rng(2).shuffle(x)
    k = np.array([1, n // 2, n // 2 + 1, n], np.uint32)
    assert np.array_equal(O.sort_choose(x, k), np.array([2, 2, 1, 1], np.float32))
    # i % 4 (generateProblems.cu:46-51): k-th largest = 3 - floor((k-1)/(n/4))
    x = (np.arange(n) % 4).astype(np.uint32)
    k = np.arange(1, n + 1, 997, dtype=np.uint32)
    assert np.array_equal(O.sort_choose(x, k), (3 - (k.astype(np.int64) - 1) // (n // 4)).astype(np.uint32))
    # duplicate-heavy uint32 [0, 1023]: the k-th largest from the histogram
    x = W.uint_range(3, n, 0, 1023)
    cnt = np.bincount(x, minlength=1024)
    above = np.cumsum(cnt[::-1])  # above[j] = #elements >= 1023 - j
    k = np.arange(1, n + 1, 101, dtype=np.uint32)
    want = 1023 - np.searchsorted(above, k.astype(np.int64))
    assert np.array_equal(O.sort_choose(x, k), want.astype(np.uint32))
    # bucket killer: all powers of two 2^-126 .. 2^126 (generateProblems.cu:363-365)
    x = (2.0 ** np.arange(-126, 127)).astype(np.float32)
    np.random.default_rng(4).shuffle(x)
    k = np.arange(1, x.size + 1, dtype=np.uint32)
    assert np.array_equal(O.sort_choose(x, k), (2.0 ** (126 - (k.astype(np.int64) - 1))).astype(np.float32))


def test_config1_golden(O, pkg):
    with open(os.path.join(ROOT, "tests", "golden", "c1_uniform_f32_2p20.json")) as f:
        g = json.load(f)
    x = pkg.workloads.uniform_f32(0, g["n"])
    assert hashlib.sha256(x.tobytes()).hexdigest() == g["input_sha256"]
    k = np.array(g["k"], np.uint32)
    assert np.array_equal(k, pkg.workloads.k_percentiles(g["n"]))
    want = np.array([int(h, 16) for h in g["expected_hex"]], np.uint32)
    assert np.array_equal(bits(O.sort_choose(x, k)), want)


@pytest.mark.parametrize("dtype", DT)
def test_topk_bruteforce(O, dtype):
    rng = np.random.default_rng(13)
    for n in (1, 2, 5, 17, 64, 1000):
        x = tiny_vectors(rng, dtype, n).astype(dtype)
        s = total_order_sort(x)[::-1]
        for K in {1, n // 2 or 1, n}:
            assert np.array_equal(bits(O.topk(x, K)), bits(s[:K]))


def test_planted_mistake_is_caught(O, pkg):
    """An off-by-one rank must change the answer on distinct values."""
    x = pkg.workloads.permutation_u32(9, 50_000).astype(np.float32)  # distinct values
    k = np.arange(2, 50_000, 37, dtype=np.uint32)
    assert not np.any(bits(O.sort_choose(x, k)) == bits(O.sort_choose(x, k - 1)))


# ------------------------------------------------------- restated algorithms
@pytest.mark.parametrize("dtype", DT)
def test_radix_select_restatement(O, dtype, pkg):
    W = pkg.workloads
    n = 30_001
    x = {np.float32: W.normal_f32(1, n), np.float64: W.cauchy_f32(2, n).astype(np.float64),
         np.uint32: W.uint_range(3, n, 0, 1023), np.int32: W.int_range(4, n, -50, 50)}[dtype]
    for kk in (1, 2, 777, n // 2, n - 1, n):
        v, dig = O.radix_select(x, kk)
        want = O.sort_choose(x, [kk])[0]
        assert bits(np.array([v], dtype))[0] == bits(np.array([want], dtype))[0]
        key = int(O.keys(np.array([want], dtype))[0])
        nd = dig.size
        nib = [(key >> (4 * (nd - 1 - i))) & 15 for i in range(nd)]
        assert list(dig) == nib


def _bucket_inputs(W, dtype, n, seed):
    if dtype == np.float32:
        return [W.normal_f32(seed, n), W.uniform_f32(seed, n), W.cauchy_f32(seed, n),
                W.plateau_f32(seed, n), (2.0 ** np.arange(-126, 127)).astype(np.float32)]
    if dtype == np.float64:
        return [W.normal_f64(seed, n), W.uniform_f64(seed, n)]
    if dtype == np.uint32:
        return [W.uint_range(seed, n, 0, 1023), W.uint_range(seed, n, 0, 2**32 - 1),
                (np.arange(n) % 4).astype(np.uint32)]
    return [W.int_range(seed, n, -2**31, 2**31 - 1), W.int_range(seed, n, -3, 3)]


@pytest.mark.parametrize("dtype", DT)
def test_bucket_multiselect_restatement(O, dtype, pkg):
    W = pkg.workloads
    for i, x in enumerate(_bucket_inputs(W, dtype, 40_000, 21)):
        n = x.size
        k = np.concatenate([W.k_uniform(30 + i, n, 300, dedup=False), [1, n, 1, n // 2]]).astype(np.uint32)
        for seed in (0, 1, 2):
            got, st = O.bucket_multiselect(x, k, seed=seed)
            assert np.array_equal(bits(got), bits(O.sort_choose(x, k)))
            assert st.total_count == n
            if not st.flat:
                assert st.monotone == 1
    # flat vector exit (bucketMultiselect.cu:541-547)
    got, st = O.bucket_multiselect(np.full(1000, 3, dtype), [1, 500, 1000])
    assert st.flat == 1 and np.all(got == 3)


@pytest.mark.parametrize("dtype", DT)
def test_randomized_topk_restatement(O, dtype, pkg):
    W = pkg.workloads
    for x in _bucket_inputs(W, dtype, 20_000, 5):
        for K in sorted({1, 10, min(257, x.size), min(4096, x.size)}):
            for seed in (0, 3):
                assert np.array_equal(bits(O.randomized_topk(x, K, seed)), bits(O.topk(x, K)))


@pytest.mark.parametrize("dtype", DT)
def test_nth_and_parallel_oracles(O, dtype, pkg):
    W = pkg.workloads
    for x in _bucket_inputs(W, dtype, 50_000, 8):
        n = x.size
        k = W.k_uniform(9, n, 500, dedup=False)
        want = bits(O.sort_choose(x, k))
        assert np.array_equal(bits(O.nth_multiselect(x, k)), want)
        assert np.array_equal(bits(O.parallel_sort_choose(x, k, threads=4)), want)
        K = min(300, n)
        assert np.array_equal(bits(O.nth_topk(x, K)), bits(O.topk(x, K)))


def test_reference_thrust_build_agrees(O, pkg):
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    x = pkg.workloads.uniform_f32(0, 1 << 20)
    k = pkg.workloads.k_percentiles(1 << 20)
    got, _ = O.ref_sort_choose(x, k, threads=4)
    assert np.array_equal(bits(got), bits(O.sort_choose(x, k)))
    for x in (pkg.workloads.normal_f64(3, 100_000), pkg.workloads.uint_range(4, 100_000, 0, 1023),
              pkg.workloads.int_range(5, 100_000, -9, 9)):
        k = pkg.workloads.k_uniform(6, x.size, 200, dedup=False)
        got, _ = O.ref_sort_choose(x, k, threads=4)
        assert np.array_equal(bits(got), bits(O.sort_choose(x, k)))
        K = 100
        got, _ = O.ref_topk(x, K, threads=4)
        assert np.array_equal(bits(got), bits(O.topk(x, K)))


def test_oracle_rejects_bad_k(O):
    x = np.arange(10, dtype=np.float32)
    for bad in ([0], [11]):
        with pytest.raises(ValueError):
            O.sort_choose(x, bad)

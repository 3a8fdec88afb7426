"""GPU regression tests for fixed defects and for the generator families.

- radix digit widths 4..7 (bucket budgets below 2^8 once left the bucket
  tables partly unloaded): full-range u32 / i32 keys with k in {1, 2, n/2,
  n}, single vector and batched;
- a batch whose engine segments hold more than 2^32 keys in total (the
  compaction once kept 32-bit positions in the reduced buffer): overlapping
  windows of one "sorted" vector, whose order statistics are single elements
  of the sequence and are computed alone on the host (no 5e9-key sort);
- process exit with live contexts, batches and servers (no destroy calls);
- every generator family: device output against the host mirror where the
  two are bit-identical, and bucket / radix multiselect against the oracle
  on the read-back data with the reference's k families;
- the Python server's tags: repeated caller tags each come back once;
- heavy values (a key repeated over 2 % .. 99.9 % of the vector: its own
  bucket) at the middle, smallest and largest keys, single and batched;
- the radix digit window below the sampled keys' common prefix, with keys
  outside it in the guard buckets;
- children of many ranks (top-K and runs of consecutive k in one bucket,
  sorted by the finisher), on both key widths.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, bits

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("rb", [4, 5, 6, 7])
def test_small_radix_digits(pkg, O, rb):
    W = pkg.workloads
    n = (1 << 20) + 7
    for dtype, x in ((np.uint32, W.uint_range(60 + rb, n, 0, 2**32 - 1)),
                     (np.int32, W.int_range(61 + rb, n, -2**31, 2**31 - 1))):
        _say(f"radix_bits={rb} {np.dtype(dtype).name} n={n}")
        # the extreme keys land in the first and last digit buckets
        x[3] = np.iinfo(dtype).max
        x[11] = np.iinfo(dtype).min
        k = np.array([1, 2, n // 2, n - 1, n], np.uint32)
        d = pkg.DeviceArray.from_host(x)
        for kind in ("radix", "bucket"):
            got = pkg.multiselect(d, k, kind=kind, radix_bits=rb)
            assert np.array_equal(bits(got), bits(O.sort_choose(x, k))), (kind, rb)
        # batched: three radix segments of different lengths back to back
        lens = [n, 70000, 300001]
        xs = [x] + [W.uint_range(62 + i, m, 0, 2**32 - 1).astype(np.uint32).view(dtype)
                    for i, m in enumerate(lens[1:])]
        xb = np.concatenate(xs)
        ks = [np.array([1, 2, m // 2, m], np.uint32) for m in lens]
        table = pkg.segment_table(lens, ["radix"] * 3, [4, 4, 4])
        b = pkg.Batch(dtype, table, np.concatenate(ks), radix_bits=rb)
        got = b(pkg.DeviceArray.from_host(xb))
        want = np.concatenate([O.sort_choose(xi, ki) for xi, ki in zip(xs, ks)])
        assert np.array_equal(bits(got), bits(want)), ("batch", rb)
        b.close()


def test_batch_beyond_2p32_keys(pkg):
    """Five engine segments of 2^30 keys each (5.4e9 in total) over one
    sorted u32 vector: segment i is the window [i 2^27, i 2^27 + 2^30), whose
    k-th largest is element i 2^27 + 2^30 - k of the sequence."""
    W = pkg.workloads
    L, step, nseg, seed = 1 << 30, 1 << 27, 5, 4242
    total = L + (nseg - 1) * step
    _say(f"generating sorted u32 of {total} keys")
    d = pkg.generate(np.uint32, "sorted", seed, total, hi=total)
    kinds = ["bucket", "topk", "radix", "bucket", "topk"]
    ks, nums = [], []
    for i, kd in enumerate(kinds):
        if kd == "topk":
            nums.append(300)
        else:
            kk = np.concatenate([[1, 2, L // 2, L - 1, L],
                                 W.k_uniform(900 + i, L, 59, dedup=False)]).astype(np.uint32)
            ks.append(kk)
            nums.append(kk.size)
    table = pkg.segment_table([L] * nseg, kinds, nums,
                              offsets=np.arange(nseg, dtype=np.uint64) * np.uint64(step))
    _say(f"batch of {nseg} x 2^30 keys, {int(table['num_k'].sum())} outputs")
    b = pkg.Batch(np.uint32, table, np.concatenate(ks))
    st = b.stats()
    assert st["engine_segments"] == nseg and st["elements"] == nseg * L
    got = b(d)
    j = 0
    for i, kd in enumerate(kinds):
        base = i * step
        if kd == "topk":
            pos = base + L - 1 - np.arange(300)
        else:
            pos = base + L - ks[j].astype(np.int64)
            j += 1
        want = np.array([W.sorted_u32(seed, 1, int(p), total)[0] for p in pos], np.uint32)
        o = int(table["out_offset"][i])
        g = got[o:o + int(table["num_k"][i])]
        bad = np.nonzero(g != want)[0]
        assert bad.size == 0, (i, kd, bad[:5], g[bad[:5]], want[bad[:5]])
    b.close()
    d.free()


def test_exit_with_live_objects():
    """A process that leaves contexts, a batch, a server and device arrays
    alive at exit (and one that destroys them) ends with status 0."""
    code = r"""
import sys, numpy as np
sys.path.insert(0, %r)
from __graft_entry__ import load_package
pkg = load_package()
W = pkg.workloads
x = W.uniform_f32(3, 1 << 20)
d = pkg.DeviceArray.from_host(x)
ctx = pkg.Context(0)
dk = pkg.DeviceArray.from_host(np.array([1, 5, 99], np.uint32))
dout = pkg.DeviceArray(3, np.float32)
pkg.multiselect_async(d, dk, dout, ctx)
pkg.multiselect(d, [1, 2, 3])          # the default context
t = pkg.segment_table([1 << 19, 1 << 19], ["bucket", "topk"], [2, 8])
b = pkg.Batch(np.float32, t, np.array([1, 7], np.uint32))
o = pkg.DeviceArray(10, np.float32)
b.run(d, o, graph=True)
srv = pkg.Server({np.float32: d}, scratch_budget=64 << 20)
out = np.zeros(4, np.float32)
srv.submit(np.float32, 0, 1 << 20, out, k=[1, 2, 3, 4])
srv.drain()
if sys.argv[1] == "destroy":
    srv.close(); b.close(); ctx.close(); d.free()
print("exit-ok", flush=True)
""" % ROOT
    for mode in ("leave", "destroy"):
        _say(f"subprocess exit test ({mode})")
        r = subprocess.run([sys.executable, "-c", code, mode], capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0 and "exit-ok" in r.stdout, (mode, r.returncode, r.stderr[-2000:])


FAMILIES_F = ["sorted", "killer", "noisy", "half_normal", "huge", "ones", "ones_twos",
              "normal100"]


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.uint32, np.int32])
def test_generator_families_and_k_families(pkg, O, dtype):
    W = pkg.workloads
    n = (1 << 20) + 3
    fams = FAMILIES_F if dtype in (np.float32, np.float64) else ["sorted", "ones", "ones_twos"]
    kfam = {"even": W.k_even(1024, n), "normal": W.k_normal(7, 1024, n),
            "cluster9": W.k_cluster9(8, 1024, n), "sectioned": W.k_sectioned(9, 1024, n),
            "uniform": W.k_uniform(10, n, 1024, dedup=False)}
    for fam in fams:
        _say(f"family {fam} {np.dtype(dtype).name}")
        d = pkg.generate(dtype, fam, 77, n, hi=n if fam == "sorted" else 0)
        x = d.to_host()
        # host mirrors that are bit-identical
        if fam == "sorted":
            assert np.array_equal(bits(x), bits(W.sorted_as(dtype, W.sorted_u32(77, n))))
            assert np.all(np.diff(x.astype(np.float64)) >= 0)
        elif fam == "killer":
            h = W.killer_f32(77, n) if dtype == np.float32 else W.killer_f64(77, n)
            assert np.array_equal(bits(x), bits(h))
        elif fam == "huge" and dtype == np.float32:
            assert np.array_equal(bits(x), bits(W.huge_f32(77, n)))
        elif fam == "ones":
            assert np.all(x == 1)
        elif fam == "ones_twos":
            assert np.array_equal(bits(x), bits(W.ones_twos(77, n, dtype)))
        elif fam == "half_normal":
            assert np.all(x >= 0)
        for kname, k in kfam.items():
            for kind in ("bucket", "radix"):
                got = pkg.multiselect(d, k, kind=kind)
                want = O.sort_choose(x, k)
                bad = np.nonzero(bits(got) != bits(want))[0]
                assert bad.size == 0, (fam, kname, kind, bad[:3])
        t = pkg.topk(d, 777)
        assert np.array_equal(bits(t), bits(O.topk(x, 777))), (fam, "topk")
        d.free()


def test_server_repeated_tags(pkg):
    W = pkg.workloads
    x = W.uniform_f32(5, 1 << 18)
    d = pkg.DeviceArray.from_host(x)
    srt = np.sort(x)
    srv = pkg.Server({np.float32: d}, scratch_budget=256 << 20, max_inflight=2, max_batch=8)
    outs = []
    for i in range(40):
        out = np.zeros(3, np.float32)
        k = np.array([1 + i, 100 + i, 1000 + i], np.uint32)
        srv.submit(np.float32, 0, x.size, out, k=k, tag=7)  # the same caller tag every time
        outs.append((k, out))
    srv.drain()
    done = srv.poll()
    assert len(done) == 40 and all(t == 7 and lat >= 0 for t, lat in done)
    for k, out in outs:
        assert np.array_equal(out, srt[x.size - k.astype(np.int64)])
    assert srv.stats()["failed"] == 0
    srv.close()


HEAVY_DT = [np.float32, np.float64, np.uint32, np.int32]


def _heavy_input(W, dtype, n, seed, frac, where):
    if dtype == np.float32:
        x = W.normal_f32(seed, n)
        v = {"mid": np.float32(0.25), "min": np.float32(-np.inf), "max": np.float32(np.inf)}[where]
    elif dtype == np.float64:
        x = W.uniform_f64(seed, n)
        v = {"mid": 0.5, "min": -np.inf, "max": np.inf}[where]
    elif dtype == np.uint32:
        x = W.uint_range(seed, n, 0, 2**32 - 1)
        v = {"mid": np.uint32(123456789), "min": np.uint32(0), "max": np.uint32(2**32 - 1)}[where]
    else:
        x = W.int_range(seed, n, -2**31, 2**31 - 1)
        v = {"mid": np.int32(-77), "min": np.int32(-2**31), "max": np.int32(2**31 - 1)}[where]
    m = int(round(frac * n))
    pos = np.argsort(W.splitmix64(seed ^ 0x5151, 0, n), kind="stable")[:m]
    x[pos] = v
    return x


@pytest.mark.parametrize("dtype", HEAVY_DT)
def test_heavy_values(pkg, O, dtype):
    """A value repeated at 2 % .. 99.9 % of the vector (its own bucket when the
    sample shows it), in the middle of the range and at the smallest and
    largest keys; ranks around it, at both ends and spread out."""
    W = pkg.workloads
    n = (1 << 22) + 5
    for frac in (0.02, 0.05, 0.5, 0.999):
        for where in ("mid", "min", "max"):
            _say(f"heavy {np.dtype(dtype).name} frac={frac} at {where}")
            x = _heavy_input(W, dtype, n, 31, frac, where)
            k = np.concatenate([W.k_uniform(5, n, 200, dedup=False), W.k_cluster(6, n, 64, 400),
                                [1, 2, n - 1, n]]).astype(np.uint32)
            want = O.sort_choose(x, k)
            d = pkg.DeviceArray.from_host(x)
            for kind in ("bucket", "radix"):
                got = pkg.multiselect(d, k, kind=kind)
                bad = np.nonzero(bits(got) != bits(want))[0]
                assert bad.size == 0, (frac, where, kind, bad[:4], got[bad[:4]], want[bad[:4]])
            t = pkg.topk(d, 300)
            assert np.array_equal(bits(t), bits(O.topk(x, 300))), (frac, where, "topk")
            d.free()


def test_heavy_values_batched(pkg, O):
    """Segments with and without a heavy value in one batch, both kinds."""
    W = pkg.workloads
    lens = [(1 << 21) + 3, 300000, 1 << 20, 77777]
    xs = [_heavy_input(W, np.float32, m, 40 + i, f, w)
          for i, (m, f, w) in enumerate(zip(lens, (0.5, 0.0, 0.1, 0.3), ("mid", "mid", "max", "min")))]
    ks = [np.concatenate([W.k_uniform(50 + i, m, 40, dedup=False), [1, m, m // 2]]).astype(np.uint32)
          for i, m in enumerate(lens)]
    for kind in ("bucket", "radix"):
        table = pkg.segment_table(lens, [kind] * 4, [k.size for k in ks])
        b = pkg.Batch(np.float32, table, np.concatenate(ks))
        got = b(pkg.DeviceArray.from_host(np.concatenate(xs)))
        want = np.concatenate([O.sort_choose(x, k) for x, k in zip(xs, ks)])
        assert np.array_equal(bits(got), bits(want)), kind
        b.close()


@pytest.mark.parametrize("dtype", [np.uint32, np.float32])
def test_radix_digit_window(pkg, O, dtype):
    """The radix digit starts below the sampled keys' common prefix: narrow
    windows anywhere in the key space, keys outside the sampled window (in
    the guard buckets), and a window that is one key wide."""
    W = pkg.workloads
    n = (1 << 21) + 9
    for lo, hi in ((0, 1023), (5000, 5000 + 70000), (2**31 - 10, 2**31 + 10),
                   (2**32 - 300, 2**32 - 1), (77, 77)):
        x = W.uint_range(lo + 1, n, lo, hi)
        # a few keys far outside the window: rarely sampled, they land in
        # the guard buckets (and are the first and last order statistics)
        x[[5, 999, 123456]] = [0, 2**32 - 1, max(lo - 1000, 0)]
        if dtype == np.float32:  # the same window of float keys
            x = (x & np.uint32(0x7F7FFFFF)).view(np.float32)
        k = np.concatenate([W.k_uniform(lo % 97, n, 100, dedup=False), [1, 2, 3, n - 1, n]]).astype(np.uint32)
        d = pkg.DeviceArray.from_host(x)
        got = pkg.multiselect(d, k, kind="radix")
        assert np.array_equal(bits(got), bits(O.sort_choose(x, k))), (lo, hi)
        d.free()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_many_rank_children(pkg, O, dtype):
    """A bucket holding hundreds of the requested ranks: top-K and runs of
    consecutive k (one run ending at the largest key) against the oracle."""
    n = 1 << 24
    _say(f"case many-rank children {np.dtype(dtype).name} n={n}")
    d = pkg.generate(dtype, "normal", 91, n)
    x = d.to_host()
    ctx = pkg.Context()
    for K in (33, 700, 4096):
        got = pkg.topk(d, K, ctx=ctx)
        want = O.sort_choose(x, np.arange(1, K + 1))
        assert np.array_equal(bits(got), bits(want)), f"top-{K}"
    runs = [np.arange(n // 3, n // 3 + 600), np.arange(1, 300), np.arange(n - 2000, n + 1)]
    for kk in runs:
        got = pkg.multiselect(d, kk, ctx=ctx)
        assert np.array_equal(bits(got), bits(O.sort_choose(x, kk))), f"run at {kk[0]}"
    ctx.close()
    d.free()

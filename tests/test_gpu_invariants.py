"""Device invariants of one engine select, read back from the context after
the call (ggks_ctx_debug_regions) and checked against what the algorithm
fixes, not against a second copy of the kernels' arithmetic:

- every key falls in exactly one bucket below B, and the reduce pass's
  bucket counts equal a host histogram of the numpy bucket function over all
  n keys (so the count kernel's one-shift fast path agrees with the plain
  bucket function), summing to n;
- the bucket function is monotone in the key over every input key and the
  edges of its clamp range and heavy value;
- a heavy-value bucket holds the heavy value alone, all of its copies;
- chunk rows are exclusive prefix sums over chunks of per-chunk histograms
  (the compaction then advances the rows of children past its LDS
  cursors by the chunk's keys of that child);
- the segment's reduced-key count (from which the compaction decides on
  child-sorted rounds and cursor windows) is the active buckets' keys;
- the children are the buckets of the active bitmap in order, packed
  contiguously, each with its exact count, a key range holding its keys
  (where one is planned), and a reduced slice equal to the multiset of its
  bucket's keys;
- each child query's rank r satisfies (keys below the bucket) + r = the
  requested order statistic's ascending rank n - k + 1, and the answers
  equal the oracle's; a child of a few ranks may be split into one-rank
  copies in the free child slots, each with its bucket's slice.

The second level is off (level2 = -1) so the child table stays as the
first level wrote it.  Cases cover the sample and radix bucket functions,
a heavy value, heavy duplication, float64 (with the LUT's sample histogram
split over work-groups) and heavy tails.
"""
import numpy as np
import pytest

from conftest import bits
from invariants import CHILD, CHUNK, QUERY, SEGDESC, SEGSTATE, bucket_of, device_counts, read, \
    regions, to_key

pytestmark = pytest.mark.gpu

CASES = [  # (label, dtype, family, lo, hi, kind, n[, number of k])
    ("f32-normal-sample", np.float32, "normal", 0, 0, "bucket", 1 << 22),
    # most keys active: the compaction sorts each round by child in LDS
    ("f32-normal-radix-dense", np.float32, "normal", 0, 0, "radix", 1 << 22),
    ("f32-plateau-heavy", np.float32, "plateau", 0, 0, "bucket", 1 << 22),
    ("f32-cauchy-sample", np.float32, "cauchy", 0, 0, "bucket", 1 << 22),
    ("u32-dup1024-sample", np.uint32, "uniform", 0, 1023, "bucket", 1 << 22),
    ("u32-dup1024-radix", np.uint32, "uniform", 0, 1023, "radix", 1 << 22),
    ("u32-full-radix", np.uint32, "uniform", 0, 0xFFFFFFFF, "radix", 1 << 22),
    ("f64-normal-sample", np.float64, "normal", 0, 0, "bucket", 1 << 21),
    # 64 Ki and 49 Ki 8-byte samples: the LUT's sample histogram is split
    # over 4 and 3 work-groups (the last one taking a remainder)
    ("f64-normal-split-lut", np.float64, "normal", 0, 0, "bucket", 1 << 24),
    ("f64-uniform-split-lut-ragged", np.float64, "uniform", 0, 0, "bucket", 50176 * 256),
    # more children than the dense sort's tables: 4-byte keys keep 4096
    # children's cursors in LDS, the rest place through the chunk rows
    ("u32-full-radix-4096k", np.uint32, "uniform", 0, 0xFFFFFFFF, "radix", 1 << 22, 4096),
    ("u32-full-radix-16384k", np.uint32, "uniform", 0, 0xFFFFFFFF, "radix", 1 << 22, 16384),
    # 8-byte keys keep 1024 cursors in LDS: the further children take
    # cursor windows (a pass over the chunk each) in the full compaction, the
    # row atomics in the small one
    ("f64-uniform-16384k-small", np.float64, "uniform", 0, 0, "bucket", 1 << 21, 16384),
    ("f64-normal-16384k", np.float64, "normal", 0, 0, "bucket", 1 << 24, 16384),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_select_invariants(pkg, O, case):
    label, dt, fam, lo, hi, kind, n = case[:7]
    nk = case[7] if len(case) > 7 else 512
    print(f"case invariants {label} n={n}", flush=True)
    lib = pkg._lib.lib()
    d = pkg.generate(dt, fam, 77, n, lo=lo, hi=hi)
    x = d.to_host()
    k = pkg.workloads.k_uniform(78, n, nk)
    ctx = pkg.Context()
    got = pkg.multiselect(d, k, ctx=ctx, kind=kind, level2=-1)
    assert np.array_equal(bits(got), bits(O.sort_choose(x, k)))
    g = regions(pkg, ctx)
    assert g.direct == 0 and g.n == n and g.nq == len(k)
    sd = read(lib, g.segs, SEGDESC, 1)[0]
    st = read(lib, g.states, SEGSTATE, 1)[0]
    B, P = int(st["B"]), int(st["P"])
    assert 2 <= B <= int(sd["bmax"]) and 1 <= P <= 1 << int(sd["pbits"])
    lut = read(lib, g.lut + 4 * int(sd["lut_off"]), np.uint32, P)

    keys = to_key(x)
    b = bucket_of(keys, st, lut)
    assert b.min() >= 0 and b.max() < B
    counts = np.bincount(b, minlength=B)
    totals = read(lib, g.totals + 4 * int(sd["tot_off"]), np.uint32, int(sd["bmax"]))
    gsums = read(lib, g.gsums + 4 * (int(sd["tot_off"]) // 64), np.uint32, int(sd["bmax"]) // 64)
    dc = device_counts(totals, gsums, B)
    assert dc.sum() == n
    bad = np.nonzero(dc != counts)[0]
    assert bad.size == 0, f"{bad.size} buckets differ, first {bad[:4]}: dev {dc[bad[:4]]} host {counts[bad[:4]]}"

    # monotone over every key and the clamp / heavy-value edges
    KT = keys.dtype.type
    W = 8 * keys.dtype.itemsize
    M = (1 << W) - 1
    edges = [0, M, int(st["kb"]), int(st["kb"]) + int(st["lmax"])]
    if int(st["hvon"]):
        edges += [int(st["hv"])]
    sweep = set()
    for e in edges:
        for dlt in (-2, -1, 0, 1, 2):
            if 0 <= e + dlt <= M:
                sweep.add(e + dlt)
    u = np.union1d(np.unique(keys), np.array(sorted(sweep), dtype=KT))
    bu = bucket_of(u, st, lut)
    assert np.all(np.diff(bu) >= 0), "bucket function not monotone"
    assert bu.max() < B
    if int(st["hvon"]):
        hv = KT(int(st["hv"]))
        hb = int(st["hb"])
        assert int(bucket_of(np.array([hv], KT), st, lut)[0]) == hb
        assert counts[hb] == int(np.count_nonzero(keys == hv)) > 0

    # chunk rows: exclusive prefix over chunks of per-chunk histograms
    nch = int(sd["nchunks"])
    chunks = read(lib, g.chunks, CHUNK, nch)
    assert np.array_equal(np.sort(chunks["idx"]), np.arange(nch))
    assert int(chunks["len"].sum()) == n
    rows = read(lib, g.rows + 4 * int(sd["row_base"]), np.uint32, nch * int(sd["bmax"]))
    rows = rows.reshape(nch, int(sd["bmax"]))
    words = read(lib, g.bits + 4 * int(sd["bits_off"]), np.uint32, (B + 31) // 32)
    active = np.nonzero(np.unpackbits(words.view(np.uint8), bitorder="little"))[0]
    # the compaction keeps the first 4096 children's cursors in LDS (1024
    # for 8-byte keys); the later children's keys are either placed through
    # their chunk's row, which so ends advanced by the chunk's keys in that
    # bucket, or (when those keys times 200 bytes outweigh re-streaming the
    # chunk) by further passes over the chunk with the next cursor windows
    # (the small compaction instance, for at most 4096 buckets and 2^10 LUT
    # prefixes, keeps 1024 cursors and no windows)
    small = (1 << int(sd["pbits"])) <= 1024 and int(sd["bmax"]) <= 4096
    curs = 4096 if x.itemsize == 4 and not small else 1024
    cnt_act = np.bincount(b, minlength=B)[active]
    beyond = int(cnt_act[curs:].sum())
    nwin = -(-active.size // curs)
    windows = not small and active.size > curs and beyond * 200 > (nwin - 1) * n * x.itemsize
    late = active[:0] if windows else active[curs:]
    print(f"children {active.size}, small {small}, beyond the LDS cursors {beyond} keys, "
          f"windows {bool(windows)}")
    run = np.zeros(B, np.int64)
    for c in chunks[np.argsort(chunks["idx"])]:
        s0 = int(c["start"])
        h = np.bincount(b[s0:s0 + int(c["len"])], minlength=B)
        want = run.copy()
        want[late] += h[late]
        assert np.array_equal(rows[int(c["idx"]), :B], want), f"row of chunk {int(c['idx'])}"
        run += h
    assert np.array_equal(run, counts)

    # children: the active buckets in order, packed, exact
    U = int(st["nact"])
    kids = read(lib, g.children + CHILD.itemsize * int(sd["q_off"]), CHILD, U)
    assert np.array_equal(kids["bucket"].astype(np.int64), active)
    assert np.all(kids["src"] == 1)
    assert np.array_equal(kids["len"].astype(np.int64), counts[active])
    ca = counts[active]
    starts = int(sd["red_base"]) + np.cumsum(ca) - ca  # exclusive prefix
    assert np.array_equal(kids["start"].astype(np.int64), starts.astype(np.int64))
    # the reduced-key count the compaction takes its dense rounds (at least a
    # quarter of the keys active, at most 1024 children) and its cursor
    # windows from
    assert int(st["nred"]) == int(ca.sum())
    if "dense" in label:
        assert U <= 1024 and 4 * int(ca.sum()) >= n
    red = read(lib, g.red + x.itemsize * int(sd["red_base"]), x.dtype, int(counts[active].sum()))
    below = np.concatenate([[0], np.cumsum(counts)])  # keys below each bucket
    order = np.argsort(b, kind="stable")  # keys grouped by bucket
    kby, xby = keys[order], bits(x)[order]
    kvals = np.asarray(k, np.int64)
    cq_all = read(lib, g.cq, QUERY, int(sd["q_off"]) + len(k))
    seen = []
    # slots past the U children: one-rank copies of split children (same
    # bucket and slice, one of its ranks each), the rest unused (no ranks)
    extra = read(lib, g.children + CHILD.itemsize * (int(sd["q_off"]) + U), CHILD, len(k) - U)
    byb = {int(c["bucket"]): c for c in kids}
    for c in extra:
        if int(c["q_cnt"]) == 0:
            continue
        assert int(c["q_cnt"]) == 1
        own = byb[int(c["bucket"])]
        assert int(c["start"]) == int(own["start"]) and int(c["len"]) == int(own["len"])
    for c in list(kids) + [c for c in extra if int(c["q_cnt"]) > 0]:
        bk = int(c["bucket"])
        sel = kby[below[bk]:below[bk + 1]]
        if not int(c["need_range"]):  # a planned key range holds the bucket's keys
            assert sel.min() >= KT(int(c["lo"])) and sel.max() <= KT(int(c["hi"])), f"range of {bk}"
        s0 = int(c["start"]) - int(sd["red_base"])
        slice_ = red[s0:s0 + int(c["len"])]
        want = np.sort(xby[below[bk]:below[bk + 1]])
        assert np.array_equal(np.sort(bits(slice_)), want), f"slice of bucket {bk}"
        for q in cq_all[int(c["q_off"]):int(c["q_off"]) + int(c["q_cnt"])]:
            r, o = int(q["rank"]), int(q["out"])
            assert 1 <= r <= int(c["len"])
            assert below[bk] + r == n - kvals[o] + 1, f"rank of query {o}"
            seen.append(o)
    # queries not sent to a child were answered from a single-key bucket
    rest = np.setdiff1d(np.arange(len(k)), np.array(seen, np.int64))
    asc = n - kvals[rest] + 1
    bq = np.searchsorted(below, asc - 1, side="right") - 1
    for j, bk in zip(rest, bq):
        sel = kby[below[bk]:below[bk + 1]]
        assert sel.size and sel.min() == sel.max(), f"query {j} answered outside a child"
    ctx.close()
    d.free()


def test_direct_select_has_no_tables(pkg):
    """A select small enough for one finisher pass plans no engine tables."""
    x = np.arange(1000, dtype=np.float32)
    d = pkg.DeviceArray.from_host(x)
    ctx = pkg.Context()
    assert np.array_equal(pkg.multiselect(d, [1, 1000], ctx=ctx), np.float32([999, 0]))
    g = regions(pkg, ctx)
    assert g.direct == 1 and g.n == 1000 and g.segs is None
    ctx.close()
    d.free()

"""CPU: the comparison driver's file formats.  Result rows in the
reference's layout (compareMultiselect.hpp:159-187) are summarised into the
reference's Summary / Overview / RatioSummary / RatioOverview files
(readMultiselectOutputfile.cpp:209-345); the expected numbers below are
worked out by hand from the rows."""
import csv
import os

import numpy as np


ROWS = [  # size, numK, arrays, ks, name, ms, name, ms, errors
    [1024, 10, "u", "k", "Sort and Choose Multiselect", 2.0, "Bucket Multiselect", 1.0, 0],
    [1024, 10, "u", "k", "Sort and Choose Multiselect", 4.0, "Bucket Multiselect", 5.0, 1],
    [1024, 20, "u", "k", "Sort and Choose Multiselect", 3.0, "Bucket Multiselect", 1.5, 0],
    [2048, 10, "u", "k", "Sort and Choose Multiselect", 8.0, "Bucket Multiselect", 2.0, 0],
]


def _write(path, rows):
    with open(path, "w", newline="") as f:
        csv.writer(f).writerows(rows)


def _read(path):
    with open(path) as f:
        return [r for r in csv.reader(f)]


def test_rows_round_trip_and_summaries(pkg, tmp_path):
    C = pkg.compare
    base = str(tmp_path / "res")
    _write(base + "_a.csv", ROWS[:2])
    _write(base + "_b.csv", ROWS[2:])
    with open(base + "_b.csv", "a") as f:  # a stray non-row line is ignored
        f.write("\n\nk value count: 10, 2^10, ratio:1\n")
    rows = C.read_rows([base + "_a.csv", base + "_b.csv"])
    assert len(rows) == 4 and rows[1][4] == [("Sort and Choose Multiselect", 4.0),
                                               ("Bucket Multiselect", 5.0)] and rows[1][5] == 1
    out = C.summarize(base)
    summ = _read(out["Summary"])
    # (1024, 10): sort mean 3, max 4, min 2, won once; bucket mean 3, max 5,
    # min 1, won once; 2 tests, 1 error
    assert [float(x) for x in summ[0]] == [1024, 10, 3, 4, 2, 1, 3, 5, 1, 1, 2, 1]
    assert [float(x) for x in summ[1]] == [1024, 20, 3, 3, 3, 0, 1.5, 1.5, 1.5, 1, 1, 0]
    over = _read(out["Overview"])
    # size 1024 over three tests: sort mean 3, bucket mean 2.5, bucket won 2
    assert [float(x) for x in over[0]] == [1024, 3, 4, 2, 1, 2.5, 5, 1, 2, 3, 1]
    rs = _read(out["RatioSummary"])
    # ratios to the first algorithm: bucket 0.5 and 1.25 -> mean 0.875
    assert rs[0][:2] == ["1024", "10"] and rs[0][6] == "Bucket Multiselect"
    assert np.isclose(float(rs[0][7]), 0.875) and np.isclose(float(rs[0][8]), 1.25)
    ro = _read(out["RatioOverview"])
    assert ro[1][0] == "2048" and np.isclose(float(ro[1][6]), 0.25)
    assert all(os.path.exists(p) for p in out.values())


def test_k_families_are_valid(pkg):
    C = pkg.compare
    W = pkg.workloads
    n = 100000
    for fam in C.K_FAMILIES:
        k = C.k_list(W, fam, 3, 500, n)
        assert k.size == 500 and k.min() >= 1 and k.max() <= n, fam
    # the reference's evenly spaced k (generateProblems.cu:506-513)
    assert list(W.k_even(4, 100)) == [1, 25, 50, 100]
    sec = W.k_sectioned(1, 50, n)
    assert np.all(np.diff(sec.astype(np.int64)) == 1)
    cl = W.k_cluster9(2, 18, n).astype(np.int64)
    assert np.all(np.diff(cl[:9]) == 1) and np.all(np.diff(cl[9:]) == 1)

"""Shared test fixtures: the package (imported by path), the oracle, markers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from __graft_entry__ import load_package  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libggks.so)")


@pytest.fixture(scope="session")
def pkg():
    return load_package()


@pytest.fixture(scope="session")
def O():
    import oracle
    oracle.build()
    return oracle


def bits(a: np.ndarray) -> np.ndarray:
    """Bit patterns, so -0.0 != +0.0 and comparisons are exact."""
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def total_order_sort(x: np.ndarray) -> np.ndarray:
    """Ascending sort in the reference's radix order, written from the order's
    definition rather than the key map: numeric order, with -0.0 before +0.0
    (and, for floats, -inf first / +inf last).  NaN-free input."""
    if x.dtype.kind == "f":
        neg_zero = (x == 0) & np.signbit(x)
        idx = np.lexsort((~neg_zero, x))
        return x[idx]
    return np.sort(x, kind="stable")

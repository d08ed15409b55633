import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def bitwise_equal(a, b):
    """Bit-for-bit equality of two float64 arrays (NaN == NaN, -0 != +0, as the reference's exact.py)."""
    import numpy as np
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def first_diff(a, b):
    import numpy as np
    ua, ub = np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64)
    idx = np.argwhere(ua != ub)
    if idx.size == 0:
        return None
    k = tuple(int(x) for x in idx[0])
    return f"{idx.shape[0]} of {a.size} differ; first at (k,j,i)={k}: {a[k]!r} vs {b[k]!r}"

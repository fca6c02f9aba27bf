import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libnspt.so)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def rel_err_per_order(a, b, nc, N):
    """max over the field of |a - b| per order / max |b| of that order.

    a, b: (nlinks, 1 + nc*nc*N) complex arrays.  Returns an array of N+1
    values (order 0 = flag)."""
    import numpy as np
    out = []
    ne = nc * nc
    for k in range(N + 1):
        sl = slice(0, 1) if k == 0 else slice(1 + (k - 1) * ne, 1 + k * ne)
        d = np.abs(a[:, sl] - b[:, sl]).max()
        s = np.abs(b[:, sl]).max()
        out.append(d / s if s > 0 else d)
    return np.array(out)


@pytest.fixture
def relerr():
    return rel_err_per_order

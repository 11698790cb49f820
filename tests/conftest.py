import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP library through the C ABI)")


def rel_err(a, b):
    """max |a-b| / max |b|"""
    import numpy as np
    a = np.asarray(a)
    b = np.asarray(b)
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    return float(np.max(np.abs(a - b))) / (scale if scale > 0 else 1.0)


@pytest.fixture
def relerr():
    return rel_err

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def rand_u64(rng, *shape):
    return rng.integers(0, 2**64, size=shape, dtype=np.uint64)


def rand_small(rng, *shape, lo=-4, hi=5):
    return rng.integers(lo, hi, size=shape).astype(np.int64).view(np.uint64)


@pytest.fixture
def rng():
    return np.random.default_rng(12345)

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "hp-multiphysics_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhpgpu.so)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def hpmp_mod():
    import hpmp
    return hpmp


def rel_err(a, b):
    """max |a - b| / max |b| over a list of array pairs (0/0 -> 0; inf when either
    side holds a NaN or an infinity, so a non-finite result never passes)."""
    num = 0.0
    den = 0.0
    for x, y in zip(a, b):
        x = np.asarray(x)
        y = np.asarray(y)
        if not (np.isfinite(x).all() and np.isfinite(y).all()):
            return float("inf")
        if x.size:
            num = max(num, float(np.abs(x - y).max()))
            den = max(den, float(np.abs(y).max()))
    return num / den if den > 0 else num

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


def shell_sources(rng, n, r0=0.8, r1=1.0, zh=0.5):
    """Uniform-in-volume hollow cylinder r in [r0, r1], |z| <= zh."""
    u = rng.random((n, 3))
    r = np.sqrt(r0 * r0 + (r1 * r1 - r0 * r0) * u[:, 0])
    ph = 2 * np.pi * u[:, 1]
    z = (2 * u[:, 2] - 1) * zh
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))

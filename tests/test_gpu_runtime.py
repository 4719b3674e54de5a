"""GPU: the library used before torch has touched the GPU leaves torch's device usable (one HIP
runtime in the process), in a fresh process."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CODE = r"""
import sys
sys.path.insert(0, '.')
import numpy as np
import bachelors_amd as fm
rng = np.random.default_rng(20261202 + 70)
F = fm.FMM(order=4, source_leaf_capacity=32, target_leaf_capacity=32)
F.set_sources(rng.uniform(-1, 1, (2000, 3)), np.full(2000, 1.0 / 2000))
F.set_targets(rng.uniform(-1, 1, (500, 3)))
B = F.run(1.0)
assert np.isfinite(B).all()
F.close()
import torch
assert torch.cuda.is_available() and torch.cuda.device_count() >= 1
assert float(torch.ones(4, device="cuda").sum()) == 4.0
print("library first, then torch: ok", flush=True)
"""


def test_library_first_then_torch():
    r = subprocess.run([sys.executable, "-c", CODE], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout[-1000:] + r.stderr[-2000:]

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


def gap_overlap_targets(orc, tgt, cap, rng, nadd=60):
    """Targets placed exactly in the one-ulp gaps and overlaps between float-rounded sibling
    boxes of the reference (bounding-box) target tree, inside the targets' bounding box so the
    root is unchanged: each lands in two children (inserted twice) or in none (dropped)."""
    T = orc.target_tree(tgt, cap)
    lo, hi = tgt.min(0), tgt.max(0)
    cand = []
    for b in np.nonzero(T.first_child >= 0)[0]:
        fc = T.first_child[b]
        for o_hi, ax in ((1, 0), (3, 1), (4, 2)):      # siblings of child 0 across x, y, z
            c_lo, c_hi, h = T.center[fc], T.center[fc + o_hi], T.half[fc]
            up, dn = c_lo[ax] + h[ax], c_hi[ax] - h[ax]
            if up != dn:
                p = c_lo.copy()
                p[ax] = min(up, dn)
                if np.all(p >= lo) and np.all(p <= hi):
                    cand.append(p)
    cand = np.array(cand)
    pick = rng.choice(len(cand), size=min(nadd, len(cand)), replace=False)
    return np.concatenate([tgt, cand[pick]]), len(pick)

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


def root_of(tree):
    """(centre, half-sizes) of an exported tree's root, from the offsets of its children."""
    c = tree["center"][0]
    kids = np.flatnonzero(tree["parent"] == 0)
    return c, 2.0 * np.abs(tree["center"][kids[0]] - c)


def coaxial_shells(rng, n_per, radii=(0.5, 0.8, 1.1, 1.4), thickness=0.1, height=1.0):
    """Coaxial cylindrical shells (uniform in volume), one random current sign per shell, equal
    weights: J = s_k phi-hat expressed as w = s_k / N."""
    P, W = [], []
    signs = rng.choice([-1.0, 1.0], size=len(radii))
    n = n_per * len(radii)
    for R, s in zip(radii, signs):
        u = rng.random((n_per, 3))
        r0, r1 = R - thickness / 2, R + thickness / 2
        r = np.sqrt(r0 * r0 + (r1 * r1 - r0 * r0) * u[:, 0])
        ph = 2 * np.pi * u[:, 1]
        z = (u[:, 2] - 0.5) * height
        P.append(np.stack([r * np.cos(ph), r * np.sin(ph), z], 1))
        W.append(np.full(n_per, s / n))
    return np.concatenate(P), np.concatenate(W)


def gaussian_clusters(rng, n, k=64):
    """n points from k anisotropic Gaussian clusters: centres U[-1,1]^3, sigma per axis U[0.01, 0.1]."""
    c = rng.uniform(-1, 1, (k, 3))
    sg = rng.uniform(0.01, 0.1, (k, 3))
    lab = rng.integers(k, size=n)
    return c[lab] + rng.normal(size=(n, 3)) * sg[lab]

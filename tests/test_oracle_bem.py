"""CPU: pins the oracle's Problem::BEM path (BEMQuadrature sources with a field B and a normal n:
the P2M of integration.cpp:295-353, calcBEM fast_multipole_solver.cpp:233-250, the root box
multipole_solver.cpp:668-669) against the direct BEM integral (integration.cpp:191-215), which is
itself checked against the closed form it restates: sum w ((y-x)(B.n) + (y-x) x (B x n)) / |y-x|^3 / 4 pi,
the gradient of the Laplace double-layer-like kernel."""
import numpy as np
import pytest

from conftest import rel_l2


def bem_sources(rng, n):
    """Points on the side of a cylinder (r = 1, |z| < 1.5) inside the reference BEM root box
    (half sizes 1.1, 1.1, 1.6), outward normals, smooth B."""
    ph = rng.uniform(0, 2 * np.pi, n)
    z = rng.uniform(-1.5, 1.5, n)
    P = np.stack([np.cos(ph), np.sin(ph), z], 1)
    N = np.stack([np.cos(ph), np.sin(ph), np.zeros(n)], 1)
    B = np.stack([0.3 * np.sin(ph), 1.0 + 0.2 * z, np.cos(2 * ph)], 1)
    w = np.full(n, 2 * np.pi * 3.0 / n)
    return P, w, B, N


def test_direct_bem_integral_closed_form(orc):
    rng = np.random.default_rng(81)
    P, w, B, N = bem_sources(rng, 200)
    x = rng.uniform(-0.5, 0.5, (20, 3))
    got = orc.direct_bem(P, w, B, N, x)
    d = P[None, :, :] - x[:, None, :]
    r3 = np.linalg.norm(d, axis=2) ** 3
    bn = np.sum(B * N, 1)
    bxn = np.cross(B, N)
    f = d * bn[None, :, None] + np.cross(d, np.broadcast_to(bxn, d.shape))
    exp = np.sum(f * (w[None, :, None] / r3[:, :, None]), 1) / (4 * 3.14159265359)
    assert np.allclose(got, exp, rtol=1e-12, atol=0)


@pytest.mark.parametrize("analytic", [False, True])
def test_bem_fmm_converges_to_the_direct_integral(orc, analytic):
    """Corrected (direct near field) BEM FMM converges to the direct integral with p; the
    reference's FD gradient in the P2M (step 1e-3) stays within its O(eps^2) of the analytic one."""
    rng = np.random.default_rng(82)
    P, w, B, N = bem_sources(rng, 6000)
    x = np.concatenate([rng.uniform(-0.8, 0.8, (300, 3)) * [1, 1, 1.5], 2.5 * rng.uniform(-1, 1, (100, 3))])
    exp = orc.direct_bem(P, w, B, N, x)
    flags = orc.CORRECTED if analytic else orc.CORRECTED & ~orc.ANALYTIC_CURL
    errs = []
    for p in (6, 10):
        R = orc.FMMRun(P, w, x, p=p, cap_s=32, cap_t=32, flags=flags, bem=(B, N))
        assert np.array_equal(np.sort(R.slots), np.arange(len(x)))
        errs.append(rel_l2(R.A_input_order(len(x)), exp))
    assert errs[1] < errs[0] / 10
    assert errs[1] < (2e-6 if analytic else 2e-5), errs

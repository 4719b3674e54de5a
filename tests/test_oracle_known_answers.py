"""CPU: the oracle's direct Biot-Savart sum against known answers -- the reference's
external Telma solution for the torus coil, and closed forms."""
import numpy as np

from torus_geometry import torus_quadratures, telma

MU0 = 1.2566370614e-6


def test_direct_sum_matches_telma(orc):
    """comparisonToTelmaIntegrals (main.cpp:50-170): direct sum on the torus coil (current 5)
    against results/telma_results.txt, to the discretisation accuracy of the 20-segment torus."""
    P, W = torus_quadratures()
    assert len(P) == 38400
    pts, Bt = telma()
    B = orc.direct_B(P, W, pts, current=5.0)
    err = np.linalg.norm(B - Bt, axis=1) / np.linalg.norm(Bt, axis=1)
    assert err.mean() < 0.02 and err.max() < 0.03
    assert np.linalg.norm(B - Bt) / np.linalg.norm(Bt) < 0.015
    # at the centre the field is axial and agrees to 1e-4
    assert abs(B[-1, 2] - Bt[-1, 2]) / Bt[-1, 2] < 1e-4
    assert np.abs(B[-1, :2]).max() < 1e-20


def test_direct_sum_circular_loop_closed_form(orc):
    """A loop of radius a carrying I: B_z(0,0,z) = mu0 I a^2 / (2 (a^2+z^2)^1.5)."""
    a, I, n = 0.9, 3.0, 4000
    ph = 2 * np.pi * (np.arange(n) + 0.5) / n
    src = np.stack([a * np.cos(ph), a * np.sin(ph), np.zeros(n)], 1)
    w = np.full(n, I * 2 * np.pi * a / n)            # line current: w = I dl, J = phi-hat
    z = np.linspace(-2, 2, 9)
    tgt = np.stack([np.zeros_like(z), np.zeros_like(z), z], 1)
    B = orc.direct_B(src, w, tgt, current=1.0)
    exact = MU0 * I * a * a / (2 * (a * a + z * z) ** 1.5)
    # the oracle keeps the reference's mu0/(4 pi) literals (math.hpp:12-13): 1e-10 relative
    assert np.allclose(B[:, 2], exact, rtol=1e-9)
    assert np.abs(B[:, :2]).max() < 1e-12 * exact.max()


def test_direct_sum_thick_solenoid_on_axis(orc):
    """Uniform azimuthal current density J in a ring a1<r<a2, |z|<=b:
    B_z(0) = mu0 J b ln((a2 + sqrt(a2^2 + b^2)) / (a1 + sqrt(a1^2 + b^2))) (both halves)."""
    rng = np.random.default_rng(3)
    a1, a2, b, J, n = 0.8, 1.0, 0.5, 2.0, 400000
    u = rng.random((n, 3))
    r = np.sqrt(a1 * a1 + (a2 * a2 - a1 * a1) * u[:, 0])
    ph = 2 * np.pi * u[:, 1]
    z = (2 * u[:, 2] - 1) * b
    src = np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)
    vol = np.pi * (a2 * a2 - a1 * a1) * 2 * b
    w = np.full(n, J * vol / n)
    B = orc.direct_B(src, w, np.zeros((1, 3)), current=1.0)
    exact = MU0 * J * b * np.log((a2 + np.hypot(a2, b)) / (a1 + np.hypot(a1, b)))
    # Monte-Carlo sampling: ~ N^-1/2
    assert abs(B[0, 2] - exact) / exact < 5e-3

"""CPU: pins the oracle's reference-faithful evaluation paths -- the finite-difference curl
(calculation_point_octree.cpp:214-275, quadrature_octree.cpp:196-258), the treecode M2P recursion
(quadrature_octree.cpp:171-258, multipole_solver.cpp:686-700) and the vector potential calcA
(fast_multipole_solver.cpp:195-212) -- against the analytic curl, the direct sums and their
convergence orders."""
import numpy as np
import pytest

from conftest import rel_l2, shell_sources


def _case(n_src=8192, n_tgt=1500, seed=41):
    rng = np.random.default_rng(seed)
    src = shell_sources(rng, n_src)
    w = rng.uniform(0.5, 1.5, n_src) / n_src
    tgt = rng.uniform(-2, 2, (n_tgt, 3))
    return src, w, tgt


def test_fd_curl_is_second_order_central_difference(orc):
    """B_far by the reference's central differences (step eps) against the analytic curl on the
    same local expansions: the gap shrinks by 4 when eps halves (a one-sided difference would give
    2, a wrong scale or sign would not shrink at all), and the reference step 1e-3 leaves it near
    1e-5 here."""
    src, w, tgt = _case()
    base = orc.SECTORAL_FIX | orc.FACT_FIX | orc.M2L_FULL | orc.P2P
    An = orc.FMMRun(src, w, tgt, p=8, cap_s=32, cap_t=32, flags=base | orc.ANALYTIC_CURL)
    gaps = []
    for eps in (2e-3, 1e-3, 5e-4):
        F = orc.FMMRun(src, w, tgt, p=8, cap_s=32, cap_t=32, flags=base, fd_eps=eps)
        assert np.array_equal(F.slots, An.slots)
        assert np.array_equal(F.Bnear, An.Bnear)          # the near field does not use the curl
        gaps.append(rel_l2(F.Bfar, An.Bfar))
    assert 3.5 < gaps[0] / gaps[1] < 4.5 and 3.5 < gaps[1] / gaps[2] < 4.5, gaps
    assert 1e-7 < gaps[1] < 1e-4


@pytest.fixture(scope="module")
def far_box_run(orc):
    """The reference's precision experiment geometry (main.cpp:962-1012): 256 targets in
    Box((10,10,10), 2), far from the sources."""
    src, w, _ = _case(n_tgt=1)
    rng = np.random.default_rng(42)
    tgt = 10.0 + rng.uniform(-2, 2, (256, 3))
    return src, w, tgt


@pytest.mark.parametrize("flags,tol", [("faithful", 1e-4), ("fixed", 1e-5)])
def test_treecode_far_points_match_direct_sums(orc, far_box_run, flags, tol):
    src, w, tgt = far_box_run
    f = orc.FAITHFUL if flags == "faithful" else orc.SECTORAL_FIX | orc.FACT_FIX
    R = orc.FMMRun(src, w, tgt, p=10, cap_s=32, cap_t=32, flags=f, current=5.0)
    B, A = R.treecode(tgt)
    assert rel_l2(B, orc.direct_B(src, w, tgt, current=5.0)) < tol
    assert rel_l2(A, orc.direct_A(src, w, tgt, current=5.0)) < tol


def test_treecode_recursion_matches_direct_sums_off_the_sources(orc):
    """Points 1 to 2.5 away from the shell: the M2P recursion opens boxes down to leaves that all
    pass 4R^2 < d^2, so it converges to the direct sums with p; closer points meet leaves that
    fail the test, which the reference drops (no P2P)."""
    src, w, _ = _case()
    rng = np.random.default_rng(43)
    ph = rng.uniform(0, 2 * np.pi, 300)
    r = rng.uniform(2.0, 3.5, 300)
    x = np.stack([r * np.cos(ph), r * np.sin(ph), rng.uniform(-0.4, 0.4, 300)], 1)
    errs = []
    for p in (6, 10):
        R = orc.FMMRun(src, w, x, p=p, cap_s=32, cap_t=32, flags=orc.SECTORAL_FIX | orc.FACT_FIX)
        B, A = R.treecode(x)
        errs.append((rel_l2(B, orc.direct_B(src, w, x)), rel_l2(A, orc.direct_A(src, w, x))))
    assert errs[1][0] < 2e-5 and errs[1][1] < 5e-6
    assert errs[1][0] < errs[0][0] / 10 and errs[1][1] < errs[0][1] / 10
    near = np.array([[0.9, 0.0, 0.0], [0.9, 0.05, 0.1]])   # among the shell's sources
    R = orc.FMMRun(src, w, near, p=10, cap_s=32, cap_t=32, flags=orc.SECTORAL_FIX | orc.FACT_FIX)
    B, _ = R.treecode(near)
    Bd = orc.direct_B(src, w, near)
    assert all(rel_l2(B[i], Bd[i]) > 1e-3 for i in range(2))


def test_fmm_far_targets_faithful_and_treecode_agree(orc, far_box_run):
    """The FMM of the reference (faithful: FD curl, l + lambda <= p M2L, sectoral quirk) on
    far-field targets has no close pairs, and agrees with the treecode and the direct sum."""
    src, w, tgt = far_box_run
    R = orc.FMMRun(src, w, tgt, p=10, cap_s=32, cap_t=32, flags=orc.FAITHFUL)
    assert R.counts["close_pairs"] == 0
    B = R.B_input_order(len(tgt))
    Bt, _ = R.treecode(tgt)
    Bd = orc.direct_B(src, w, tgt)
    assert rel_l2(B, Bd) < 1e-4 and rel_l2(Bt, Bd) < 1e-4
    assert rel_l2(B, Bt) < 1e-4


def test_calcA_converges_to_the_direct_vector_potential(orc):
    src, w, tgt = _case(n_tgt=1000)
    Ad = orc.direct_A(src, w, tgt, current=3.0)
    errs = []
    for p in (6, 10):
        R = orc.FMMRun(src, w, tgt, p=p, cap_s=32, cap_t=32, flags=orc.CORRECTED, current=3.0)
        A = R.A_input_order(len(tgt))
        assert np.allclose(R.A, R.Afar + R.Anear, rtol=1e-12, atol=0)
        errs.append(rel_l2(A, Ad))
    assert errs[1] < 1e-6 and errs[1] < errs[0] / 10
    # faithful: the close part by M2P only (close leaves dropped) is far off, as the reference
    F = orc.FMMRun(src, w, tgt, p=10, cap_s=32, cap_t=32, flags=orc.FAITHFUL, current=3.0)
    assert rel_l2(F.A_input_order(len(tgt)), Ad) > 1e-3

"""GPU: every M2M algorithm of the reference (M2MAlg, multipole_solver.cpp:102-339) gives the
multipoles and the field of the Matrices path: NoTranslation (P2M at every node from its points),
ComplexTranslation / RealTranslation / Layers (the per-node translation formula, no operator
tables) and Matrices (the MFMA GEMM of per-(level, orientation) operators).  In F16X3 the P2M
also splits its leaves' M2L columns and the split after the M2M covers the nodes the algorithm
made, so a wrong split of either part shows up in B."""
import numpy as np
import pytest

from conftest import rel_l2, shell_sources

pytestmark = pytest.mark.gpu


def _run(fm, src, w, tgt, precision, alg):
    F = fm.FMM(order=8, source_leaf_capacity=24, target_leaf_capacity=24, precision=precision, frame=fm.FRAME_SHARED)
    F.set_sources(src, w)
    F.set_targets(tgt)
    F.build_trees()
    F.leaf_expansions()
    F.upward(alg)
    F.downward()
    B = F.evaluate_B(1.0)
    S = F.tree(0)
    nodes = np.flatnonzero(S["first_child"] >= 0)
    M = np.stack([F.expansion(0, n) for n in nodes])
    return B, M


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("f16x3", 2e-6)])
def test_every_m2m_algorithm_gives_the_matrices_result(precision, tol):
    import bachelors_amd as fm
    rng = np.random.default_rng(71)
    src = shell_sources(rng, 5000)
    w = rng.uniform(0.5, 1.5, len(src)) / len(src)
    tgt = rng.uniform(-2, 2, (1200, 3))
    B0, M0 = _run(fm, src, w, tgt, precision, fm.M2MAlg.Matrices)
    assert len(M0) > 50 and np.abs(M0).max() > 0
    for alg in (fm.M2MAlg.NoTranslation, fm.M2MAlg.ComplexTranslation, fm.M2MAlg.RealTranslation,
                fm.M2MAlg.Layers):
        B, M = _run(fm, src, w, tgt, precision, alg)
        # multipoles of every subdivided node (the leaves are the same P2M in every algorithm)
        for k in range(len(M)):
            assert np.abs(M[k] - M0[k]).max() <= tol * np.abs(M0[k]).max(), (alg, k)
        assert rel_l2(B, B0) < tol, alg

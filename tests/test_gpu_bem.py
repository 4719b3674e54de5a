"""GPU: Problem::BEM (BEMQuadrature sources with a field B and a normal; P2M of the gradient of the
regular harmonics, calcBEM, the reference's BEM root box) against the oracle and the direct BEM
integral (integration.cpp:191-215)."""
import numpy as np
import pytest

from conftest import rel_l2
from test_oracle_bem import bem_sources

pytestmark = pytest.mark.gpu


def _targets(rng):
    inner = rng.uniform(-0.8, 0.8, (300, 3)) * [1, 1, 1.5]
    return np.concatenate([inner, 2.5 * rng.uniform(-1, 1, (100, 3))])


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 3e-6)])
def test_bem_matches_oracle_and_direct_integral(orc, precision, tol):
    import bachelors_amd as fm
    rng = np.random.default_rng(91)
    P, w, B, N = bem_sources(rng, 6000)
    x = _targets(rng)
    s = fm.FastMultipoleSolver(np.column_stack([P, w]), x, problem=fm.Problem.BEM, quadratureOctreeLeafCapacity=32,
                               calculationPointOctreeLeafCapacity=32, order=10, precision=precision, bem_B=B,
                               bem_normal=N)
    s.calcMultipoleExpansionsAtLeaves()
    s.calclMultipoleExpansions(fm.M2MAlg.Matrices, fm.Device.GPU)
    s.calcLocalMultipoleExpansions(fm.M2LAlg.ComplexTranslation)
    V = s.calcBEM_array(5.0)
    R = orc.FMMRun(P, w, x, p=10, cap_s=32, cap_t=32, flags=orc.CORRECTED, bem=(B, N))
    Vo = R.A_input_order(len(x))
    ok = np.isfinite(Vo[:, 0])
    assert np.array_equal(np.isfinite(V[:, 0]), ok) and ok.sum() >= len(x) - 2
    assert rel_l2(V[ok], Vo[ok]) < tol
    assert rel_l2(V[ok], orc.direct_bem(P, w, B, N, x[ok])) < 2e-6
    pairs = s.calcBEM(5.0)
    assert len(pairs) == ok.sum()


def test_bem_reference_mode(orc):
    """The reference BEM path: close lists by M2P walks (failing leaves dropped), M2L l + lambda <= p;
    against the oracle on the same rules with the exact P2M gradient, and the stated gap to the
    reference as written (central-difference gradient, step 1e-3, sectoral quirk)."""
    import bachelors_amd as fm
    rng = np.random.default_rng(92)
    P, w, B, N = bem_sources(rng, 6000)
    x = _targets(rng)
    F = fm.FMM(order=10, source_leaf_capacity=32, target_leaf_capacity=32, precision="fp64", m2l_full=0,
               near_field=fm.NEAR_TREECODE, source_root_center=(0, 0, 0), source_root_half=(1.1, 1.1, 1.6))
    F.set_bem_sources(P, w, B, N)
    F.set_targets(x)
    F.build_trees()
    F.leaf_expansions()
    F.upward()
    F.downward()
    V = F.evaluate_bem()
    same = orc.FMMRun(P, w, x, p=10, cap_s=32, cap_t=32, flags=orc.SECTORAL_FIX | orc.FACT_FIX | orc.ANALYTIC_CURL,
                      bem=(B, N))
    Vo = same.A_input_order(len(x))
    ok = np.isfinite(Vo[:, 0])
    assert rel_l2(V[ok], Vo[ok]) < 1e-10
    ref = orc.FMMRun(P, w, x, p=10, cap_s=32, cap_t=32, flags=orc.FAITHFUL, bem=(B, N))
    gap = rel_l2(V[ok], ref.A_input_order(len(x))[ok])
    print("rel-L2 to the reference BEM FMM as written: %.3e" % gap)
    assert gap < 1e-4


def test_bem_shared_frame_split_precisions(orc):
    """BEM sources in the shared frame (class-batched M2L) with a cubic source root of half 2 around
    the reference's 1.1 x 1.1 x 1.6 box (with that box, or a cube of half 1.6, the library refuses
    the shared frame: interaction offsets off the box lattice), in FP32 and F16X3 against FP64 on
    the same trees, and FP64 against the direct BEM integral (measured: 5.3e-8, 5.2e-8, 2.1e-7).
    BEM leaves keep the whole-array clear and the split of every node's M2L columns after the M2M
    (their P2M does not split)."""
    import bachelors_amd as fm
    rng = np.random.default_rng(93)
    P, w, B, N = bem_sources(rng, 6000)
    x = _targets(rng)

    def run(precision):
        F = fm.FMM(order=10, source_leaf_capacity=32, target_leaf_capacity=32, precision=precision,
                   frame=fm.FRAME_SHARED, source_root_half=(2.0, 2.0, 2.0))
        F.set_bem_sources(P, w, B, N)
        F.set_targets(x)
        F.build_trees()
        F.leaf_expansions()
        F.upward()
        F.downward()
        V = F.evaluate_bem()
        far = F.stats()["far_pairs"]
        F.close()
        return V, far

    V64, far = run("fp64")
    ok = np.isfinite(V64[:, 0])
    assert far > 1000 and ok.sum() >= len(x) - 2
    e64 = rel_l2(V64[ok], orc.direct_bem(P, w, B, N, x[ok]))
    print("fp64 vs direct %.3e, far pairs %d" % (e64, far), flush=True)
    assert e64 < 2e-6
    for precision in ("fp32", "f16x3"):
        V, _ = run(precision)
        e = rel_l2(V[ok], V64[ok])
        print("%s vs fp64 %.3e" % (precision, e), flush=True)
        assert np.array_equal(np.isfinite(V[:, 0]), ok)
        assert e < 2e-6, precision

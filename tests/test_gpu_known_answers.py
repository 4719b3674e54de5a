"""GPU: the FMM on the reference's torus coil (Telma known answer).

The Telma points reach outside the coil's bounding box, so the shared target root is grown
from the source root by doublings: the case where source and target tree levels differ in
box size and M2L classes must be keyed by physical box size."""
import numpy as np
import pytest

from conftest import rel_l2, root_of
from torus_geometry import torus_quadratures, telma

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 3e-6)])
def test_fmm_torus_against_oracle_direct_and_telma(orc, precision, tol):
    import bachelors_amd as fm
    P, W = torus_quadratures()
    pts, Bt = telma()
    # the reference's own settings of comparisonToTelmaIntegrals: capS 1000, capT 32 (main.cpp:72)
    s = fm.FastMultipoleSolver(np.column_stack([P, W]), pts, quadratureOctreeLeafCapacity=1000,
                               calculationPointOctreeLeafCapacity=32, order=10, precision=precision,
                               frame=fm.FRAME_SHARED)
    s.calcMultipoleExpansionsAtLeaves()
    s.calclMultipoleExpansions(fm.M2MAlg.Matrices, fm.Device.GPU)
    s.calcLocalMultipoleExpansions(fm.M2LAlg.ComplexTranslation, fm.Device.CPU)
    B = s.calcB_array(5.0)
    T = s._fmm.tree(1)
    root = root_of(T)
    assert np.all(root[1] > 0)
    # parity: the oracle's FMM on the same frame, order and leaf capacities
    R = orc.FMMRun(P, W, pts, p=10, cap_s=1000, cap_t=32, flags=orc.CORRECTED, target_root=root, current=5.0)
    assert rel_l2(B, R.B_input_order(len(pts))) < tol
    # truncation at p = 10 (measured 1.2e-6 against the direct sum)
    Bd = orc.direct_B(P, W, pts, current=5.0)
    assert rel_l2(B, Bd) < 5e-6
    assert rel_l2(B, Bt) < 0.015
    # reference output order: pairs in target-leaf order, each point once
    pairs = s.calcB(5.0)
    assert len(pairs) == len(pts)
    assert sorted(map(tuple, (p for p, _ in pairs))) == sorted(map(tuple, pts))

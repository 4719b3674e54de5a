"""GPU: the FMM on inputs shaped like the BASELINE configurations, at sizes the oracle finishes in
seconds -- config 1 in full (the CPU reference case), config 3 (4 coaxial shells with random-sign
currents, p=16, split precision) and config 5 (anisotropic Gaussian clusters with random current
directions, p=14, leaf capacity 48, depth <= 10) reduced.  Parity is against the oracle's FMM at the
same order and tree (same frame, target root taken from the GPU tree); truncation against the direct
sum with the bound measured for the oracle itself."""
import numpy as np
import pytest

from conftest import coaxial_shells, gaussian_clusters, rel_l2, root_of, shell_sources

pytestmark = pytest.mark.gpu


def _gpu(fm, src, w, tgt, J=None, **cfg):
    F = fm.FMM(**cfg)
    F.set_sources(src, w, J)
    F.set_targets(tgt)
    return F, F.run(1.0)


@pytest.mark.parametrize("frame", [0, 1])
def test_config1_reference_case_fp64(orc, frame):
    import bachelors_amd as fm
    rng = np.random.default_rng(20261202 + 1)
    src = shell_sources(rng, 32768)
    w = np.full(32768, 1.0 / 32768)
    tgt = rng.uniform(-2, 2, (4096, 3))
    F, B = _gpu(fm, src, w, tgt, order=10, source_leaf_capacity=32, target_leaf_capacity=32, precision="fp64",
                frame=frame)
    root = root_of(F.tree(1)) if frame == 1 else None
    R = orc.FMMRun(src, w, tgt, p=10, cap_s=32, cap_t=32, flags=orc.CORRECTED, target_root=root)
    assert rel_l2(B, R.B_input_order(len(tgt))) < 1e-10
    # p = 10 truncation of this geometry: 1.55e-6 for the oracle itself
    assert rel_l2(B, orc.direct_B(src, w, tgt)) < 3e-6


def test_config3_coaxial_shells_p16(orc):
    import bachelors_amd as fm
    rng = np.random.default_rng(20261202 + 3)
    src, w = coaxial_shells(rng, 1024)
    tgt = rng.uniform(-2, 2, (2048, 3))
    cfg = dict(order=16, source_leaf_capacity=64, target_leaf_capacity=64, frame=1)
    F, B64 = _gpu(fm, src, w, tgt, precision="fp64", **cfg)
    R = orc.FMMRun(src, w, tgt, p=16, cap_s=64, cap_t=64, flags=orc.CORRECTED, target_root=root_of(F.tree(1)))
    assert rel_l2(B64, R.B_input_order(len(tgt))) < 1e-10
    assert rel_l2(B64, orc.direct_B(src, w, tgt)) < 1e-8        # oracle's own truncation: 2.98e-9
    _, B32 = _gpu(fm, src, w, tgt, precision="fp32", **cfg)
    _, Bbf = _gpu(fm, src, w, tgt, precision="bf16x3", **cfg)
    _, B16 = _gpu(fm, src, w, tgt, precision="f16x3", **cfg)
    assert rel_l2(B32, B64) < 1e-6      # measured 7.2e-8
    assert rel_l2(Bbf, B64) < 2e-6      # measured 2.1e-7
    assert rel_l2(B16, B64) < 1e-6      # measured 7.1e-8


def test_config5_gaussian_clusters_p14(orc):
    import bachelors_amd as fm
    rng = np.random.default_rng(20261202 + 5)
    src = gaussian_clusters(rng, 2048)
    J = rng.normal(size=(2048, 3))
    J /= np.linalg.norm(J, axis=1)[:, None]
    w = np.full(2048, 1.0 / 2048)
    tgt = gaussian_clusters(np.random.default_rng(20261202 + 50), 2048)
    cfg = dict(order=14, source_leaf_capacity=48, target_leaf_capacity=48, frame=1, max_depth=10)
    F, B64 = _gpu(fm, src, w, tgt, J=J, precision="fp64", **cfg)
    R = orc.FMMRun(src, w, tgt, p=14, cap_s=48, cap_t=48, flags=orc.CORRECTED, J=J, target_root=root_of(F.tree(1)),
                   max_depth=10)
    assert rel_l2(B64, R.B_input_order(len(tgt))) < 1e-10
    assert rel_l2(B64, orc.direct_B(src, w, tgt, J=J)) < 3e-8   # oracle's own truncation: 6.6e-9
    _, B32 = _gpu(fm, src, w, tgt, J=J, precision="fp32", **cfg)
    _, Bbf = _gpu(fm, src, w, tgt, J=J, precision="bf16x3", **cfg)
    _, B16 = _gpu(fm, src, w, tgt, J=J, precision="f16x3", **cfg)
    assert rel_l2(B32, B64) < 1e-6      # measured 1.7e-7
    assert rel_l2(Bbf, B64) < 2e-6      # measured 2.9e-7
    assert rel_l2(B16, B64) < 1e-6      # measured 1.7e-7


@pytest.mark.parametrize("order", [13, 14, 15, 16, 17, 18])
def test_f16x3_orders_13_to_18_match_fp64(order):
    """The F16X3 M2L instances of the high orders against the FP64 FMM on the same trees: p = 13
    (single-buffered kernel), 15 and 18 (persistent kernel, K padded to 32), 14 and 16 (persistent,
    the last coefficient as an FP32 rank-1 update in the epilogue) and 17 (the last four)."""
    import bachelors_amd as fm
    rng = np.random.default_rng(20261202 + 60 + order)
    src, w = coaxial_shells(rng, 4096)
    tgt = rng.uniform(-2, 2, (4096, 3))
    cfg = dict(order=order, source_leaf_capacity=32, target_leaf_capacity=32, frame=1)
    _, B64 = _gpu(fm, src, w, tgt, precision="fp64", **cfg)
    _, B16 = _gpu(fm, src, w, tgt, precision="f16x3", **cfg)
    assert np.isfinite(B16).all()
    assert rel_l2(B16, B64) < 1e-6      # measured 6.3e-8 .. 9.3e-8 over the six orders


@pytest.mark.parametrize("ns,nt", [(40, 10), (300, 120), (3000, 500)])
def test_f16x3_p16_fused_small_inputs(orc, ns, nt):
    """The fused evaluation at p = 16 in F16X3 (persistent M2L GEMM, near field beside it) on inputs
    so small that there are no far pairs at all (40 sources: one leaf, the near field only) or
    only a few tiles, against the FP64 direct sum and against the staged evaluation."""
    import ctypes
    import torch
    import bachelors_amd as fm
    rng = np.random.default_rng(20261202 + 80 + ns)
    src, w = coaxial_shells(rng, ns // 4)
    tgt = rng.uniform(-2, 2, (nt, 3))
    F = fm.FMM(order=16, source_leaf_capacity=64, target_leaf_capacity=64, precision="f16x3", frame=1)
    F.set_sources(src, w)
    F.set_targets(tgt)
    B_staged = F.run(1.0)
    out = torch.empty((nt, 3), dtype=torch.float64, device="cuda")
    F.run_B_dev(1.0, ctypes.c_void_p(out.data_ptr()))
    F.synchronize()
    B = out.cpu().numpy()
    assert np.isfinite(B).all()
    assert rel_l2(B, B_staged) < 1e-6
    assert rel_l2(B, orc.direct_B(src, w, tgt)) < 1e-6
    if ns == 40:
        assert F.stats()["far_pairs"] == 0

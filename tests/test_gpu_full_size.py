"""GPU, full size: the workloads the throughput is quoted on, checked through properties that do not
need the CPU oracle -- rel-L2 against the library's FP64 direct sum on the first 16,384 targets,
the fused evaluation equal to the staged one, the adjoint equal to the transpose (g.B(w) = w.B^T g,
B linear in w) -- plus the reduced-size parity against the oracle at the same order and leaf
capacity:
  * 1,048,576 shell sources, 1,048,576 uniform targets, p = 12, leaf capacity 64 (FP32, F16X3);
  * 4,194,304 sources in 4 coaxial shells, 4,194,304 uniform targets, p = 16 (F16X3, FP64);
  * the adjoint at p = 12: 1,048,576 shell sources with weights U[0.5, 1.5], 262,144 targets;
  * 8,388,608 sources and targets from 64 anisotropic Gaussian clusters with random current
    directions, p = 14, leaf capacity 48, depth <= 10: forward and backward."""
import ctypes

import numpy as np
import pytest

from conftest import coaxial_shells, gaussian_clusters, rel_l2, root_of, shell_sources

pytestmark = pytest.mark.gpu

SUBSET = 16384


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _direct(fm, d_src, d_w, d_tgt, d_J=None, m=SUBSET):
    import torch
    D = fm.FMM(order=2, source_leaf_capacity=64, target_leaf_capacity=64)
    D.set_sources_dev(d_src.shape[0], _vp(d_src), _vp(d_w), None if d_J is None else _vp(d_J))
    D.set_targets_dev(m, _vp(d_tgt))
    D.build_trees()
    out = torch.empty((m, 3), dtype=torch.float64, device="cuda")
    D.direct_B_dev(1.0, m, _vp(d_tgt), _vp(out))
    D.close()
    return out.cpu().numpy()


def _fused(fm, d_src, d_w, d_tgt, d_J=None, **cfg):
    import torch
    F = fm.FMM(frame=fm.FRAME_SHARED, **cfg)
    F.set_sources_dev(d_src.shape[0], _vp(d_src), _vp(d_w), None if d_J is None else _vp(d_J))
    F.set_targets_dev(d_tgt.shape[0], _vp(d_tgt))
    B = torch.empty_like(d_tgt)
    F.run_B_dev(1.0, _vp(B))
    F.synchronize()
    return F, B


@pytest.fixture(scope="module")
def shell1m():
    rng = np.random.default_rng(20261202 + 2)
    src = shell_sources(rng, 1 << 20)
    w = np.full(1 << 20, 1.0 / (1 << 20))
    tgt = rng.uniform(-2, 2, (1 << 20, 3))
    return src, w, tgt


@pytest.mark.parametrize("precision", ["fp32", "f16x3"])
def test_shell_1m_p12(shell1m, precision):
    import torch
    import bachelors_amd as fm
    src, w, tgt = shell1m
    d_src, d_w, d_tgt = _dev(src), _dev(w), _dev(tgt)
    ref = _direct(fm, d_src, d_w, d_tgt)
    F, B = _fused(fm, d_src, d_w, d_tgt, order=12, source_leaf_capacity=64, target_leaf_capacity=64,
                  precision=precision)
    Bh = B.cpu().numpy()
    assert np.isfinite(Bh).all()
    assert rel_l2(Bh[:SUBSET], ref) < 1e-6          # measured 1.5e-7
    # the stage-by-stage calls give the same field (the fused one runs lists and P2P beside)
    out = torch.empty_like(d_tgt)
    F.build_trees(); F.leaf_expansions(); F.upward(); F.downward(); F.evaluate_B_dev(1.0, _vp(out))
    F.synchronize()
    assert rel_l2(out.cpu().numpy(), Bh) < 1e-6
    F.close()


def test_shell_reduced_p12_cap64_matches_oracle(orc):
    """The same geometry at 32,768 sources and 8,192 targets, p = 12, leaf capacity 64: the GPU
    against the oracle on the same trees (shared frame), FP32 and F16X3 within FP32 rounding."""
    import bachelors_amd as fm
    rng = np.random.default_rng(20261202 + 20)
    src = shell_sources(rng, 32768)
    w = np.full(32768, 1.0 / 32768)
    tgt = rng.uniform(-2, 2, (8192, 3))
    R = None
    for precision, tol in (("fp64", 1e-10), ("fp32", 1e-6), ("f16x3", 1e-6)):
        F = fm.FMM(order=12, source_leaf_capacity=64, target_leaf_capacity=64, precision=precision,
                   frame=fm.FRAME_SHARED)
        F.set_sources(src, w)
        F.set_targets(tgt)
        B = F.run(1.0)
        if R is None:
            R = orc.FMMRun(src, w, tgt, p=12, cap_s=64, cap_t=64, flags=orc.CORRECTED, target_root=root_of(F.tree(1)))
            Bo = R.B_input_order(len(tgt))
        assert rel_l2(B, Bo) < tol, precision
    assert rel_l2(Bo, orc.direct_B(src, w, tgt)) < 1e-6


def test_coaxial_shells_4m_p16():
    import bachelors_amd as fm
    rng = np.random.default_rng(20261202 + 3)
    src, w = coaxial_shells(rng, 1 << 20)
    tgt = rng.uniform(-2, 2, (1 << 22, 3))
    d_src, d_w, d_tgt = _dev(src), _dev(w), _dev(tgt)
    ref = _direct(fm, d_src, d_w, d_tgt)
    errs = {}
    for precision in ("f16x3", "fp64"):
        F, B = _fused(fm, d_src, d_w, d_tgt, order=16, source_leaf_capacity=64, target_leaf_capacity=64,
                      precision=precision)
        Bh = B[:SUBSET].cpu().numpy()
        errs[precision] = rel_l2(Bh, ref)
        F.close()
    print("4M coaxial shells, p=16: rel-L2 vs FP64 direct", errs)
    assert errs["fp64"] < 1e-8
    assert errs["f16x3"] < 1e-6


def _transpose_gap(fm, src, w, tgt, J=None, precision="fp64", **cfg):
    """|g.B(w) - w.B^T g| / |g.B(w)| through the C ABI (device pointers), for g = B(w) times
    1 + U[-1/2, 1/2] per component (fixed seed): a generic direction whose product with B(w) is
    at least |B(w)|^2 / 2, so the quotient measures the FMM error and not the cancellation of a
    random-sign sum."""
    import torch
    d_src, d_w, d_tgt = _dev(src), _dev(w), _dev(tgt)
    d_J = None if J is None else _dev(J)
    F, B = _fused(fm, d_src, d_w, d_tgt, d_J, precision=precision, **cfg)
    u = np.random.default_rng(20261202 + 7).uniform(-0.5, 0.5, tuple(B.shape))
    g = B * (1 + _dev(u))
    gw = torch.empty_like(d_w)
    F.adjoint_weights_dev(1.0, _vp(g), _vp(gw))
    F.synchronize()
    lhs = float((B * g).sum())
    rhs = float((d_w * gw).sum())
    F.close()
    return abs(lhs - rhs) / abs(lhs), gw.cpu().numpy()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("fp32", 2e-6), ("f16x3", 2e-6)])
def test_adjoint_1m_p12_is_the_transpose(precision, tol):
    import bachelors_amd as fm
    rng = np.random.default_rng(20261202 + 4)
    src = shell_sources(rng, 1 << 20)
    w = rng.uniform(0.5, 1.5, 1 << 20)
    tgt = rng.uniform(-2, 2, (1 << 18, 3))
    gap, _ = _transpose_gap(fm, src, w, tgt, precision=precision, order=12, source_leaf_capacity=64,
                            target_leaf_capacity=64)
    assert gap < tol


def test_clusters_8m_p14_forward_and_backward():
    """8,388,608 clustered sources with random current directions and as many clustered targets,
    p = 14, leaf capacity 48, depth <= 10 (the capped leaves at the cluster centres hold about a
    thousand points): F16X3 forward within 1e-6 of the FP64 direct sum, F16X3 and FP64 backward
    equal to their transposes, and the two gradients within FP32 rounding of each other."""
    import bachelors_amd as fm
    rng = np.random.default_rng(20261202 + 5)
    n = 1 << 23
    src = gaussian_clusters(rng, n)
    J = rng.normal(size=(n, 3))
    J /= np.linalg.norm(J, axis=1)[:, None]
    w = np.full(n, 1.0 / n)
    tgt = gaussian_clusters(np.random.default_rng(20261202 + 50), n)
    cfg = dict(order=14, source_leaf_capacity=48, target_leaf_capacity=48, max_depth=10)
    d_src, d_w, d_tgt, d_J = _dev(src), _dev(w), _dev(tgt), _dev(J)
    ref = _direct(fm, d_src, d_w, d_tgt, d_J)
    F, B = _fused(fm, d_src, d_w, d_tgt, d_J, precision="f16x3", **cfg)
    st = F.stats()
    assert st["source_levels"] == 11 and st["target_levels"] == 11     # the depth cap is reached
    assert rel_l2(B[:SUBSET].cpu().numpy(), ref) < 1e-6
    F.close()
    del d_src, d_w, d_tgt, d_J, B
    gaps, grads = {}, {}
    for precision in ("f16x3", "fp64"):
        gaps[precision], grads[precision] = _transpose_gap(fm, src, w, tgt, J, precision=precision, **cfg)
    assert gaps["fp64"] < 1e-11 and gaps["f16x3"] < 2e-6
    assert rel_l2(grads["f16x3"], grads["fp64"]) < 1e-5


def test_clusters_depth_capped_adjoint_vs_oracle_gradient(orc):
    """Clusters tight enough (sigma 1e-3) that 4,096 sources reach the depth cap 10 with heavy
    leaves: the F16X3 and BF16X3 adjoints with random current directions against the FP64 direct
    gradient of the same loss (oracle), and against the FP64 adjoint."""
    import torch
    from bachelors_amd.torch_fmm import FMMBiotSavart
    rng = np.random.default_rng(20261202 + 55)
    c = rng.uniform(-1, 1, (8, 3))
    lab = rng.integers(8, size=4096)
    src = c[lab] + rng.normal(size=(4096, 3)) * 1e-3
    J = rng.normal(size=(4096, 3))
    J /= np.linalg.norm(J, axis=1)[:, None]
    w = rng.uniform(0.5, 1.5, 4096) / 4096
    tgt = c[rng.integers(8, size=1024)] + rng.normal(size=(1024, 3)) * 0.05
    wstar = w * (1 + 0.1 * (2 * rng.random(len(w)) - 1))
    grads = {}
    for precision in ("fp64", "f16x3", "bf16x3"):
        F = FMMBiotSavart(src, tgt, J=J, order=14, cap=48, precision=precision, max_depth=10)
        with torch.no_grad():
            Bstar = F(torch.tensor(wstar, device="cuda"))
        wt = torch.tensor(w, dtype=torch.float64, device="cuda", requires_grad=True)
        ((F(wt) - Bstar) ** 2).sum().backward()
        grads[precision] = wt.grad.cpu().numpy()
        if precision == "fp64":
            assert F._fmm.stats()["source_levels"] == 11
    assert rel_l2(grads["f16x3"], grads["fp64"]) < 1e-5
    assert rel_l2(grads["bf16x3"], grads["fp64"]) < 1e-4
    Bs = orc.direct_B(src, wstar, tgt, J=J)
    gd = orc.direct_grad_w(src, w, tgt, Bs, J=J)
    assert rel_l2(grads["fp64"], gd) < 1e-5

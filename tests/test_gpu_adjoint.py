"""GPU: the hand-written adjoint FMM (dL/dw for L = sum_i |B_i - B*_i|^2) through the C ABI
and the torch.autograd harness."""
import numpy as np
import pytest

from conftest import rel_l2, shell_sources

pytestmark = pytest.mark.gpu


def _case(seed=21, ns=6000, nt=1500):
    rng = np.random.default_rng(seed)
    src = shell_sources(rng, ns)
    w = rng.uniform(0.5, 1.5, ns)
    tgt = rng.uniform(-2, 2, (nt, 3))
    return rng, src, w, tgt


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("fp32", 2e-6), ("f16x3", 2e-6)])
def test_adjoint_is_the_transpose(precision, tol):
    """B is linear in w, so g.B(w) == w.(adjoint g) for every g and w (exact up to rounding,
    independent of the FMM truncation error)."""
    import torch
    from bachelors_amd.torch_fmm import FMMBiotSavart
    rng, src, w, tgt = _case()
    F = FMMBiotSavart(src, tgt, order=8, cap=32, precision=precision)
    wt = torch.tensor(w, dtype=torch.float64, device="cuda", requires_grad=True)
    B = F(wt)
    g = torch.tensor(rng.normal(size=(len(tgt), 3)), dtype=torch.float64, device="cuda")
    (B * g).sum().backward()
    lhs = float((B * g).sum())
    rhs = float((wt.detach() * wt.grad).sum())
    assert abs(lhs - rhs) <= tol * abs(lhs)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("fp32", 2e-6), ("f16x3", 2e-6)])
def test_adjoint_is_the_transpose_p16(precision, tol):
    """The same identity at p = 16 (the order of the 4M workload; in F16X3 the forward runs the
    persistent M2L GEMM with the near field beside it), with g = B(w) (1 + U[-1/2, 1/2]) so that
    g.B(w) >= |B(w)|^2 / 2 and the quotient is not a random-sign cancellation."""
    import torch
    from bachelors_amd.torch_fmm import FMMBiotSavart
    rng, src, w, tgt = _case(23, ns=8000, nt=2000)
    F = FMMBiotSavart(src, tgt, order=16, cap=64, precision=precision)
    wt = torch.tensor(w, dtype=torch.float64, device="cuda", requires_grad=True)
    B = F(wt)
    g = B.detach() * (1 + torch.tensor(rng.uniform(-0.5, 0.5, (len(tgt), 3)), device="cuda"))
    (B * g).sum().backward()
    lhs = float((B * g).sum())
    rhs = float((wt.detach() * wt.grad).sum())
    assert abs(lhs - rhs) <= tol * abs(lhs)


def test_adjoint_gradient_vs_finite_differences_and_direct():
    import torch
    from bachelors_amd.torch_fmm import FMMBiotSavart
    from oracle import oracle as O
    rng, src, w, tgt = _case(22)
    wstar = w * (1 + 0.1 * (2 * rng.random(len(w)) - 1))
    F = FMMBiotSavart(src, tgt, order=10, cap=32, precision="fp64")
    with torch.no_grad():
        Bstar = F(torch.tensor(wstar, device="cuda"))
    wt = torch.tensor(w, dtype=torch.float64, device="cuda", requires_grad=True)
    loss = ((F(wt) - Bstar) ** 2).sum()
    loss.backward()
    grad = wt.grad.cpu().numpy()
    # central finite differences, h = 1e-3, on 64 seeded weights
    idx = rng.choice(len(w), 64, replace=False)
    fd = []
    with torch.no_grad():
        for j in idx:
            wp = w.copy(); wp[j] += 1e-3
            wm = w.copy(); wm[j] -= 1e-3
            lp = float(((F(torch.tensor(wp, device="cuda")) - Bstar) ** 2).sum())
            lm = float(((F(torch.tensor(wm, device="cuda")) - Bstar) ** 2).sum())
            fd.append((lp - lm) / 2e-3)
    fd = np.array(fd)
    assert rel_l2(grad[idx], fd) < 1e-5
    # the adjoint operator alone: the FMM residual r = B(w) - B* pulled back through the exact
    # transposed kernel, dL/dw_j = 2 sum_i r_i . (MU0/4pi) (J_j x (x_i - y_j)) / |x_i - y_j|^3
    with torch.no_grad():
        r = (F(torch.tensor(w, device="cuda")) - Bstar).cpu().numpy()
    J = np.stack([-src[:, 1], src[:, 0], np.zeros(len(src))], 1) / np.hypot(src[:, 0], src[:, 1])[:, None]
    gk = np.zeros(len(src))
    for j0 in range(0, len(src), 500):
        d = tgt[None, :, :] - src[j0:j0 + 500, None, :]                        # (j, i, 3)
        k = np.cross(np.broadcast_to(J[j0:j0 + 500, None, :], d.shape), d) / (np.linalg.norm(d, axis=2) ** 3)[..., None]
        gk[j0:j0 + 500] = 2 * np.einsum("jic,ic->j", k, r) * (1.2566370614e-6 / (4 * 3.14159265359))
    assert rel_l2(grad, gk) < 1e-5      # p = 10 truncation of the transposed FMM
    # against the FP64 direct-sum gradient of the same loss (direct-sum B and B*): the FMM error of
    # the residual (a tenth of B) enters here too
    Bs = O.direct_B(src, wstar, tgt)
    gd = O.direct_grad_w(src, w, tgt, Bs)
    assert rel_l2(grad, gd) < 1e-4

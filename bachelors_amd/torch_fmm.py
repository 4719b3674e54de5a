"""torch.autograd harness for the FMM: B(w) forward, transposed FMM backward.

FMMBiotSavart(src, tgt, ...) holds one FMM context whose trees, lists and operator
tables depend only on the positions; calling it with per-source weights w (a CUDA
float64 tensor) returns B at the targets (input order), and backward gives dL/dw through
the library's hand-written adjoint (fmm_adjoint_weights_dev), not through autograd of
the kernels.  This is a thin test harness over the C ABI.
"""
import ctypes

import torch

import bachelors_amd as fm


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


class _Apply(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w, module):
        w = w.detach().contiguous()
        module._fmm.set_weights_dev(_vp(w))
        module._fmm.leaf_expansions()
        module._fmm.upward()
        module._fmm.downward()
        B = torch.empty((module.nt, 3), dtype=torch.float64, device=w.device)
        module._fmm.evaluate_B_dev(module.current, _vp(B))
        module._fmm.synchronize()
        ctx.module = module
        return B

    @staticmethod
    def backward(ctx, gB):
        m = ctx.module
        gB = gB.detach().contiguous().to(torch.float64)
        gw = torch.empty(m.ns, dtype=torch.float64, device=gB.device)
        m._fmm.adjoint_weights_dev(m.current, _vp(gB), _vp(gw))
        m._fmm.synchronize()
        return gw, None


class FMMBiotSavart:
    def __init__(self, src, tgt, J=None, order=12, cap=64, precision="fp64", current=1.0, w0=None, **config):
        """config: further fmm_config fields (max_depth, mac_scale, m2l_full, ...)."""
        dev = torch.device("cuda:0")
        self.src = torch.as_tensor(src, dtype=torch.float64, device=dev).contiguous()
        self.tgt = torch.as_tensor(tgt, dtype=torch.float64, device=dev).contiguous()
        self.J = None if J is None else torch.as_tensor(J, dtype=torch.float64, device=dev).contiguous()
        self.ns, self.nt = self.src.shape[0], self.tgt.shape[0]
        self.current = current
        w = torch.ones(self.ns, dtype=torch.float64, device=dev) if w0 is None else \
            torch.as_tensor(w0, dtype=torch.float64, device=dev).contiguous()
        self._fmm = fm.FMM(order=order, source_leaf_capacity=cap, target_leaf_capacity=cap, precision=precision,
                           frame=fm.FRAME_SHARED, **config)
        self._fmm.set_sources_dev(self.ns, _vp(self.src), _vp(w), None if self.J is None else _vp(self.J))
        self._fmm.set_targets_dev(self.nt, _vp(self.tgt))
        self._fmm.build_trees()

    def __call__(self, w):
        return _Apply.apply(w, self)

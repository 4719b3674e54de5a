"""GPU: fmm_direct_B_noalloc_dev (the allocation-free direct sum, device scratch array) against
fmm_direct_B_dev (library workspace), through the C ABI: two calls in a row, the second smaller,
then a call above the size limit that must be refused without running."""
import ctypes

import numpy as np
import pytest

from conftest import shell_sources

pytestmark = pytest.mark.gpu


def test_noalloc_direct_sum_matches_the_allocating_one():
    import torch
    import bachelors_amd as fm
    L = fm.lib()
    L.fmm_direct_B_noalloc_dev.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_size_t, ctypes.c_void_p,
                                           ctypes.c_void_p]
    rng = np.random.default_rng(61)
    src = shell_sources(rng, 50000)
    w = rng.uniform(0.5, 1.5, len(src)) / len(src)
    tgt = rng.uniform(-2, 2, (16384, 3))
    F = fm.FMM(order=4, source_leaf_capacity=64, target_leaf_capacity=64)
    F.set_sources(src, w)
    F.set_targets(tgt[:10])
    F.build_trees()
    t = torch.from_numpy(tgt).cuda()
    for m in (16384, 3001):     # the second call smaller: stale scratch from the first must not leak in
        ref = torch.full((m, 3), np.nan, dtype=torch.float64, device="cuda")
        out = torch.full((m, 3), np.nan, dtype=torch.float64, device="cuda")
        F.direct_B_dev(2.0, m, ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(ref.data_ptr()))
        rc = L.fmm_direct_B_noalloc_dev(F.h, 2.0, m, ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(out.data_ptr()))
        assert rc == 0, L.fmm_last_error(F.h)
        torch.cuda.synchronize()
        a, b = out.cpu().numpy(), ref.cpu().numpy()
        assert np.isfinite(a).all()
        # same source ranges in both entry points at these sizes: the same sums, bit for bit
        assert np.array_equal(a, b)
    # above the limit: refused before any GPU work, with a message
    big = fm.FMM(order=4, source_leaf_capacity=64, target_leaf_capacity=64)
    big.set_sources(src[:100], w[:100])
    big.set_targets(tgt[:10])
    big.build_trees()
    rc = L.fmm_direct_B_noalloc_dev(big.h, 1.0, 65537, ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(t.data_ptr()))
    assert rc == 4 and b"65536" in L.fmm_last_error(big.h)

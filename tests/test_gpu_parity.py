"""GPU parity: the HIP library (through its C ABI) against the CPU oracle on the same
seeded inputs.  Integer work (trees, lists) must match bit for bit; floating point to
the tolerance written in each test."""
import numpy as np
import pytest

from conftest import gap_overlap_targets, rel_l2, shell_sources

pytestmark = pytest.mark.gpu

SHARED_ROOT = (np.zeros(3), np.full(3, 3.0))


@pytest.fixture(scope="module")
def fm():
    import bachelors_amd
    bachelors_amd.lib()
    return bachelors_amd


def _inputs(seed, ns, nt, kind="shell"):
    rng = np.random.default_rng(seed)
    src = shell_sources(rng, ns)
    w = np.full(ns, 1.0 / ns)
    if kind == "uniform":
        tgt = rng.uniform(-2, 2, (nt, 3))
    else:
        tgt = np.concatenate([rng.uniform(-2, 2, (nt // 2, 3)), shell_sources(rng, nt - nt // 2)])
    return src, w, tgt


def _boundary_targets(orc, tgt, cap, rng, nadd=400):
    """Targets placed exactly on child-box faces of the reference (bounding-box) target
    tree: exercises half-open containment and the float-rounded one-ulp gaps/overlaps."""
    T = orc.target_tree(tgt, cap)
    sub = np.nonzero(T.first_child >= 0)[0]
    pts = []
    for _ in range(nadd):
        b = sub[rng.integers(len(sub))]
        ch = T.first_child[b] + rng.integers(8)
        c, h = T.center[ch], T.half[ch]
        p = c + rng.uniform(-1, 1, 3) * h
        ax = rng.integers(3)
        p[ax] = c[ax] + (h[ax] if rng.random() < 0.5 else -h[ax])
        pts.append(p)
    return np.concatenate([tgt, np.array(pts)])


def _cmp_tree(g, o):
    assert len(g["level"]) == len(o.level)
    for k in ("level", "parent", "first_child", "begin", "end", "npts"):
        assert np.array_equal(g[k], getattr(o, k)), k
    assert np.array_equal(g["center"], o.center)
    assert np.array_equal(g["slots"], o.slots)


@pytest.mark.parametrize("frame", [0, 1])
def test_trees_bit_exact(fm, orc, frame):
    rng = np.random.default_rng(7)
    src, w, tgt = _inputs(3, 20000, 6000)
    if frame == 0:
        tgt = _boundary_targets(orc, tgt, 24, rng)
    F = fm.FMM(order=4, source_leaf_capacity=20, target_leaf_capacity=24, frame=frame)
    F.set_sources(src, w)
    F.set_targets(tgt)
    F.build_trees()
    _cmp_tree(F.tree(0), orc.source_tree(src, 20))
    _cmp_tree(F.tree(1), orc.target_tree(tgt, 24, root=SHARED_ROOT if frame == 1 else None))
    assert F.stats()["source_tree_sorted"] == 1


def test_trees_bit_exact_float_gaps_and_overlaps(fm, orc):
    """Reference frame: targets inside one-ulp gaps/overlaps of float-rounded sibling boxes are
    dropped or inserted twice, as by the reference's pointer insertion (level-by-level build)."""
    rng = np.random.default_rng(23)
    src, w, tgt = _inputs(3, 8000, 6000)
    tgt, nadd = gap_overlap_targets(orc, tgt, 24, rng)
    F = fm.FMM(order=4, source_leaf_capacity=20, target_leaf_capacity=24, frame=0)
    F.set_sources(src, w)
    F.set_targets(tgt)
    F.build_trees()
    _cmp_tree(F.tree(1), orc.target_tree(tgt, 24))
    assert F.stats()["target_tree_sorted"] == 0


def test_trees_bit_exact_lattice_faces_and_deep_cluster(fm, orc):
    """Shared frame: targets exactly on child faces (the half-open lattice boxes tile exactly,
    so the one-sort build applies) and a source cluster deeper than 10 levels (63-bit paths)."""
    rng = np.random.default_rng(17)
    src, w, tgt = _inputs(5, 12000, 3000)
    src = np.concatenate([src, 0.3 + rng.uniform(-1e-5, 1e-5, (3000, 3))])
    w = np.full(len(src), 1.0 / len(src))
    h = 3.0 / 2 ** rng.integers(1, 7, 500)            # faces of the shared lattice, levels 1..6
    faces = rng.uniform(-2, 2, (500, 3))
    ax = rng.integers(3, size=500)
    faces[np.arange(500), ax] = np.round(faces[np.arange(500), ax] / h) * h
    tgt = np.concatenate([tgt, faces])
    F = fm.FMM(order=4, source_leaf_capacity=20, target_leaf_capacity=24, frame=1)
    F.set_sources(src, w)
    F.set_targets(tgt)
    F.build_trees()
    S = orc.source_tree(src, 20)
    assert S.level.max() > 10
    _cmp_tree(F.tree(0), S)
    _cmp_tree(F.tree(1), orc.target_tree(tgt, 24, root=SHARED_ROOT))
    st = F.stats()
    assert st["source_tree_sorted"] == 1 and st["target_tree_sorted"] == 1


@pytest.mark.parametrize("max_depth", [1, 3, 5])
def test_trees_bit_exact_depth_capped(fm, orc, max_depth):
    """A depth cap leaves nodes above the leaf capacity unsplit."""
    src, w, tgt = _inputs(11, 20000, 6000)
    F = fm.FMM(order=4, source_leaf_capacity=20, target_leaf_capacity=24, frame=1, max_depth=max_depth)
    F.set_sources(src, w)
    F.set_targets(tgt)
    F.build_trees()
    S = orc.source_tree(src, 20, max_depth=max_depth)
    assert S.level.max() == max_depth
    _cmp_tree(F.tree(0), S)
    _cmp_tree(F.tree(1), orc.target_tree(tgt, 24, root=SHARED_ROOT, max_depth=max_depth))


@pytest.mark.parametrize("frame", [0, 1])
def test_interaction_lists_bit_exact(fm, orc, frame):
    src, w, tgt = _inputs(4, 12000, 4000)
    F = fm.FMM(order=4, source_leaf_capacity=24, target_leaf_capacity=24, frame=frame)
    F.set_sources(src, w)
    F.set_targets(tgt)
    F.build_trees()
    S = orc.source_tree(src, 24)
    T = orc.target_tree(tgt, 24, root=SHARED_ROOT if frame == 1 else None)
    (oft, ofs), (oct_, ocs) = orc.interaction_lists(S, T)
    F.leaf_expansions()
    F.upward()
    F.downward()
    (ft, fs), (ct, cs), ncls = F.lists()
    assert (ncls > 0) == (frame == 1)
    assert set(zip(ft.tolist(), fs.tolist())) == set(zip(oft.tolist(), ofs.tolist()))
    assert set(zip(ct.tolist(), cs.tolist())) == set(zip(oct_.tolist(), ocs.tolist()))
    assert len(ft) == len(oft) and len(ct) == len(oct_)


def _class_codes(S, T, ft, fs):
    """(level difference, centre offset in units of the finer box half size) of each far pair,
    from the oracle's trees, as one sortable integer."""
    hf = np.minimum(S.half[fs, 0], T.half[ft, 0])
    off = np.rint((S.center[fs] - T.center[ft]) / hf[:, None]).astype(np.int64) + 128
    d = (S.level[fs] - T.level[ft]).astype(np.int64) + 32
    return (d << 24) | (off[:, 0] << 16) | (off[:, 1] << 8) | off[:, 2]


@pytest.mark.parametrize("near_uniform", [True, False])
def test_far_pairs_class_order(fm, orc, near_uniform):
    """Shared frame: far pairs come sorted by (class, target) and the class count matches the
    distinct (level difference, offset) codes.  Uniform targets as dense as the sources give
    level differences up to 2 and offsets up to 15 finer half sizes; few targets around a
    thin shell give level differences up to 4 and offsets up to 51."""
    if near_uniform:
        src, w, tgt = _inputs(6, 16000, 16000, kind="uniform")
        caps = capt = 32
    else:
        src, w, tgt = _inputs(6, 12000, 4000)
        caps, capt = 20, 24
    F = fm.FMM(order=4, source_leaf_capacity=caps, target_leaf_capacity=capt, frame=1)
    F.set_sources(src, w)
    F.set_targets(tgt)
    F.build_trees()
    S = orc.source_tree(src, caps)
    T = orc.target_tree(tgt, capt, root=SHARED_ROOT)
    (oft, ofs), _ = orc.interaction_lists(S, T)
    F.leaf_expansions()
    F.upward()
    F.downward()
    (ft, fs), _, ncls = F.lists()
    code = _class_codes(S, T, oft, ofs)
    order = np.lexsort((oft, code))
    assert np.array_equal(ft, oft[order]) and np.array_equal(fs, ofs[order])
    assert ncls == len(np.unique(code))
    d = (code >> 24) - 32
    o = np.stack([(code >> 16) & 255, (code >> 8) & 255, code & 255], 1) - 128
    assert (d.max(), np.abs(o).max()) == ((2, 15) if near_uniform else (4, 51))


@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("p", [4, 12])
def test_operator_tables_fp64(fm, orc, kind, p):
    rng = np.random.default_rng(kind * 10 + p)
    deg = np.array([int(np.sqrt(i)) for i in range((p + 1) ** 2)])
    for _ in range(3):
        if kind == 1:
            t = rng.uniform(-3, 3, 3) + np.array([4.0, 0, 0])
            si, so = 0.5, 1.0
        else:
            t = rng.uniform(-0.5, 0.5, 3)
            si, so = (0.25, 0.5) if kind == 0 else (0.5, 0.25)
        G = fm.make_operator(kind, p, t, si, so, 1)
        T = orc.operator(kind, p, t, orc.CORRECTED)
        if kind == 0:
            E = T * (si ** deg[None, :]) / (so ** deg[:, None])
        elif kind == 2:
            E = T * (so ** (deg[:, None] + 1.0)) / (si ** (deg[None, :] + 1.0))
        else:
            E = T * (so ** (deg[:, None] + 1.0)) * (si ** deg[None, :])
        # 1e-10: the oracle keeps the reference's truncated 1/sqrt2 (math.hpp:16)
        assert np.abs(G - E).max() <= 1e-10 * np.abs(E).max()


def _oracle_run(orc, src, w, tgt, p, caps, cap_t, flags=None):
    return orc.FMMRun(src, w, tgt, p=p, cap_s=caps, cap_t=cap_t, flags=orc.CORRECTED if flags is None else flags,
                      target_root=SHARED_ROOT)


def _gpu_run(fm, src, w, tgt, p, caps, cap_t, precision):
    F = fm.FMM(order=p, source_leaf_capacity=caps, target_leaf_capacity=cap_t, precision=precision, frame=1)
    F.set_sources(src, w)
    F.set_targets(tgt)
    B = F.run(1.0)
    return F, B


@pytest.fixture(scope="module")
def small_case(orc):
    src, w, tgt = _inputs(11, 4096, 1500)
    R = _oracle_run(orc, src, w, tgt, 8, 32, 32)
    return src, w, tgt, R


def test_expansions_fp64(fm, orc, small_case):
    src, w, tgt, R = small_case
    F, _ = _gpu_run(fm, src, w, tgt, 8, 32, 32, "fp64")
    S = orc.source_tree(src, 32)
    T = orc.target_tree(tgt, 32, root=SHARED_ROOT)
    checked = 0
    for b in range(0, len(S.level), max(1, len(S.level) // 40)):
        if S.first_child[b] < 0 and S.npts[b] == 0:
            continue
        Mo = R.multipole(b)
        Mg = F.expansion(0, b)
        assert np.abs(Mg - Mo).max() <= 1e-9 * max(np.abs(Mo).max(), 1e-300)
        checked += 1
    for b in range(0, len(T.level), max(1, len(T.level) // 40)):
        if T.first_child[b] < 0 and T.npts[b] == 0:
            continue
        Lo = R.local(b)
        Lg = F.expansion(1, b)
        if np.abs(Lo).max() == 0:
            assert np.abs(Lg).max() == 0
        else:
            assert np.abs(Lg - Lo).max() <= 1e-9 * np.abs(Lo).max()
        checked += 1
    assert checked > 20


# measured: fp32 2.2e-7, bf16x3 9.6e-7, f16x3 2.1e-7
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-6), ("bf16x3", 1e-5), ("f16x3", 1e-6)])
def test_B_matches_oracle(fm, orc, small_case, precision, tol):
    src, w, tgt, R = small_case
    F, B = _gpu_run(fm, src, w, tgt, 8, 32, 32, precision)
    Bo = R.B_input_order(len(tgt))
    assert np.all(np.isfinite(B))
    assert rel_l2(B, Bo) < tol
    Bf, Bn = F.B_parts()
    assert rel_l2(Bn, R.Bnear) < (1e-12 if precision == "fp64" else 1e-6)
    assert rel_l2(Bf, R.Bfar) < tol * 10


def test_direct_sum_gpu_vs_oracle(fm, orc):
    import ctypes
    src, w, tgt = _inputs(12, 3000, 500)
    F = fm.FMM(order=4, source_leaf_capacity=32, target_leaf_capacity=32)
    F.set_sources(src, w)
    F.set_targets(tgt)
    F.build_trees()
    Bo = orc.direct_B(src, w, tgt, current=2.0)
    import torch
    t = torch.from_numpy(tgt).cuda()
    out = torch.zeros_like(t)
    F.direct_B_dev(2.0, len(tgt), ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(out.data_ptr()))
    assert rel_l2(out.cpu().numpy(), Bo) < 1e-13


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 3e-6)])
def test_reference_frame_fmm_matches_oracle(fm, orc, precision, tol):
    """The whole FMM on the reference's own trees (bounding-box target root, float-rounded
    centres, per-pair M2L/L2L operators) against the oracle on the same trees."""
    src, w, tgt = _inputs(13, 3000, 800)
    F = fm.FMM(order=8, source_leaf_capacity=32, target_leaf_capacity=32, precision=precision, frame=0)
    F.set_sources(src, w)
    F.set_targets(tgt)
    B = F.run(1.0)
    R = orc.FMMRun(src, w, tgt, p=8, cap_s=32, cap_t=32, flags=orc.CORRECTED)
    Bf, Bn = F.B_parts()
    assert rel_l2(Bf + Bn, R.Bfar + R.Bnear) < tol
    Bo = R.B_input_order(len(tgt))
    ok = np.isfinite(Bo[:, 0])
    assert np.array_equal(np.isfinite(B[:, 0]), ok)
    assert rel_l2(B[ok], Bo[ok]) < tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32", 1e-6), ("f16x3", 1e-6)])
def test_fused_run_matches_staged_run(fm, precision, tol):
    """fmm_run_B_dev (lists and near field on a side stream) equals the stage-by-stage calls; the
    only difference allowed is the order of the atomic M2M additions on narrow levels."""
    import torch
    src, w, tgt = _inputs(21, 20000, 6000)
    F = fm.FMM(order=8, source_leaf_capacity=32, target_leaf_capacity=32, precision=precision, frame=1)
    F.set_sources(src, w)
    F.set_targets(tgt)
    B_staged = F.run(1.0)
    out = torch.empty((len(tgt), 3), dtype=torch.float64, device="cuda")
    for _ in range(2):
        F.run_B_dev(1.0, out.data_ptr())
        F.synchronize()
        assert rel_l2(out.cpu().numpy(), B_staged) < tol

"""CPU: pins the oracle's trees (quadrature_octree.cpp / calculation_point_octree.cpp
restated) and interaction maps (fast_multipole_solver.cpp:105-184) by brute force."""
import numpy as np
import pytest

from conftest import gap_overlap_targets, shell_sources

CHILD_SIGN = np.array([(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1),
                       (-1, 1, 1)])


def _in_box(P, c, h):
    return np.all((c - h <= P) & (P < c + h), axis=1)


def _adversarial(rng, n):
    """Points on dyadic planes of the [-3,3)^3 root, plus random ones."""
    k = rng.integers(0, 5, (n, 3))
    P = -3 + 6 * rng.integers(0, 2 ** k) / 2.0 ** k
    mix = rng.random((n, 3)) < 0.5
    return np.where(mix, P, rng.uniform(-3, 3, (n, 3)))


@pytest.mark.parametrize("kind", ["shell", "adversarial", "clustered"])
def test_source_tree_rules(orc, kind):
    rng = np.random.default_rng(11)
    if kind == "shell":
        P = shell_sources(rng, 3000)
    elif kind == "adversarial":
        P = _adversarial(rng, 3000)
        P[:5] = 3.0          # outside the half-open root: dropped (quadrature_octree.cpp:33-36)
    else:
        P = rng.normal(scale=0.05, size=(3000, 3)) + 0.3
    cap = 16
    T = orc.source_tree(P, cap)
    inside = _in_box(P, np.zeros(3), np.full(3, 3.0))
    # every point inside the root lands in exactly one leaf, once
    assert sorted(T.slots.tolist()) == np.nonzero(inside)[0].tolist()
    for b in range(len(T.level)):
        cnt = int(_in_box(P, T.center[b], T.half[b]).sum())
        sub = T.first_child[b] >= 0
        # subdivided <=> population >= cap (source leaf holds <= cap-1, quadrature_octree.cpp:38);
        # coincident points stop at the 21-level depth cap (the reference recurses without bound)
        assert sub == (cnt >= cap and T.level[b] < 21), (b, cnt, sub)
        assert T.end[b] - T.begin[b] == cnt
        if not sub:
            idx = T.slots[T.begin[b]:T.end[b]]
            assert np.all(np.diff(idx) > 0)                 # insertion (input) order
            assert np.all(_in_box(P[idx], T.center[b], T.half[b]))
        else:
            fc = T.first_child[b]
            for o in range(8):
                ch = fc + o
                assert T.parent[ch] == b and T.level[ch] == T.level[b] + 1
                # reference child order and half sizes (quadrature_octree.cpp:88-96)
                assert np.array_equal(np.sign(T.center[ch] - T.center[b]), CHILD_SIGN[o])
                assert np.allclose(T.half[ch], T.half[b] / 2, rtol=0, atol=0)


def test_target_tree_capacity_and_bbox_root(orc):
    rng = np.random.default_rng(12)
    P = rng.uniform(-2, 2, (2500, 3))
    cap = 20
    T = orc.target_tree(P, cap)
    c, h = orc.bounding_box(P)
    # math.cpp:39-57
    mn, mx = P.min(0), P.max(0)
    assert np.array_equal(c, mn + ((mx - mn) / 2 + 1e-5))
    assert np.array_equal(h, ((mx - mn) / 2 + 1e-5) * 1.1)
    assert np.array_equal(T.center[0], c) and np.array_equal(T.half[0], h)
    for b in range(len(T.level)):
        if T.first_child[b] < 0:
            assert T.npts[b] <= cap           # target leaf holds <= cap (calculation_point_octree.cpp:41)
        else:
            # children centres rounded through float (calculation_point_octree.cpp:76-96)
            f = np.float32
            fc = T.first_child[b]
            for o in range(8):
                exp = np.array([f(f(T.center[b][a]) + f(CHILD_SIGN[o][a]) * (f(T.half[b][a]) / f(2)))
                                for a in range(3)], dtype=np.float64)
                assert np.array_equal(T.center[fc + o], exp)
    # population rule on nodes whose boxes tile their parent exactly is subdivided <=> count > cap
    assert T.first_child[0] >= 0


def _target_rule_violations(T, P, cap, max_depth=21):
    """Nodes of a target tree on a dyadic root (children tile their parent exactly, so a node's
    arrivals are the points in its box) that break calculation_point_octree.cpp:41: a node
    subdivides exactly when its population reaches cap + 1 (a leaf holds at most cap)."""
    bad = []
    for b in range(len(T.level)):
        cnt = int(_in_box(P, T.center[b], T.half[b]).sum())
        if (T.first_child[b] >= 0) != (cnt >= cap + 1 and T.level[b] < max_depth) or T.end[b] - T.begin[b] != cnt:
            bad.append((b, cnt))
    return bad


@pytest.mark.parametrize("kind", ["uniform", "adversarial"])
def test_target_tree_threshold_by_brute_force(orc, kind):
    """Target tree on the dyadic root [-3, 3)^3: subdivided <=> population >= cap + 1, checked by
    recounting every node's box; the data holds nodes of exactly cap and cap + 1 points, so an
    off-by-one threshold cannot pass (planted below)."""
    rng = np.random.default_rng(31)
    P = rng.uniform(-2, 2, (4000, 3)) if kind == "uniform" else _adversarial(rng, 4000)
    cap = 12
    root = (np.zeros(3), np.full(3, 3.0))
    T = orc.target_tree(P, cap, root=root)
    assert sorted(T.slots.tolist()) == list(range(len(P)))
    assert _target_rule_violations(T, P, cap) == []
    pops = [int(_in_box(P, T.center[b], T.half[b]).sum()) for b in range(len(T.level))]
    leaf_pops = {c for b, c in enumerate(pops) if T.first_child[b] < 0}
    sub_pops = {c for b, c in enumerate(pops) if T.first_child[b] >= 0}
    assert cap in leaf_pops and cap + 1 in sub_pops        # both sides of the threshold occur
    # planted mistake: the source tree's threshold (subdivide at cap) on the same points
    wrong = orc.Tree(P, cap, root=root)
    assert _target_rule_violations(wrong, P, cap) != []


def _leaf_of(T, n):
    leaf = np.full(n, -1)
    for b in range(len(T.level)):
        if T.first_child[b] < 0:
            leaf[T.slots[T.begin[b]:T.end[b]]] = b
    return leaf


def _ancestors(T, b):
    out = []
    while b >= 0:
        out.append(b)
        b = T.parent[b]
    return out


@pytest.mark.parametrize("shared", [False, True])
def test_interaction_lists_cover_every_pair_once(orc, shared):
    rng = np.random.default_rng(13)
    src = shell_sources(rng, 1500)
    tgt = rng.uniform(-2, 2, (400, 3))
    S = orc.source_tree(src, 12)
    root = (np.zeros(3), np.full(3, 3.0)) if shared else None
    T = orc.target_tree(tgt, 10, root=root)
    (ft, fs), (ct, cs) = orc.interaction_lists(S, T)
    rad_s = np.linalg.norm(S.half, axis=1)
    rad_t = np.linalg.norm(T.half, axis=1)

    def far(s, t):
        md = 2 * max(rad_s[s], rad_t[t]) + min(rad_s[s], rad_t[t])
        d = S.center[s] - T.center[t]
        return md * md < d[0] * d[0] + d[1] * d[1] + d[2] * d[2]

    assert all(far(s, t) for t, s in zip(ft, fs))
    assert not any(far(s, t) for t, s in zip(ct, cs))
    # close pairs only at useful target leaves
    assert np.all(T.npts[ct] > 0)
    # coverage: every (target point, source point) pair counted exactly once
    src_leaf = _leaf_of(S, len(src))
    far_set = {}
    for t, s in zip(ft, fs):
        far_set.setdefault(int(t), []).append(int(s))
    close_set = {}
    for t, s in zip(ct, cs):
        close_set.setdefault(int(t), []).append(int(s))
    src_anc = [set(_ancestors(S, src_leaf[j])) for j in range(len(src))]
    for b in range(len(T.level)):
        if T.npts[b] == 0:
            continue
        hits = np.zeros(len(src), int)
        for a in _ancestors(T, b):
            for s in far_set.get(a, []):
                hits += np.array([s in anc for anc in src_anc])
        for s in close_set.get(b, []):
            hits += np.array([s in anc for anc in src_anc])
        assert np.all(hits == 1), (b, np.unique(hits))


def test_gap_overlap_targets_are_duplicated_or_dropped(orc):
    """The construction used by the GPU fallback test really hits float gaps/overlaps: the
    reference tree holds some of those points twice or not at all."""
    rng = np.random.default_rng(23)
    tgt = np.concatenate([rng.uniform(-2, 2, (3000, 3)), shell_sources(rng, 3000)])
    aug, nadd = gap_overlap_targets(orc, tgt, 24, rng)
    assert nadd >= 20
    T = orc.target_tree(aug, 24)
    assert np.array_equal(T.center[0], orc.target_tree(tgt, 24).center[0])   # same root
    counts = np.bincount(T.slots, minlength=len(aug))
    extra = counts[len(tgt):]
    assert np.any(extra != 1)


def _rule_violations(S, T, ft, fs, ct, cs, mac_scale=1.0):
    """Checks every far and close pair against the descent rule of formInteractionMaps
    (fast_multipole_solver.cpp:105-165) bottom-up: the pair's own fate is far / close, and some
    chain of generating pairs leads back to the root pair, each link being a pair whose fate under
    the rule is to descend into the side that produced the next link.  The rule at a pair (t, s)
    that is neither far nor close: descend into s's children when (R_t < R_s or t is a useful
    leaf) and s is not a useful leaf, else into t's children; at a useful t it descends into s when
    some child of s is far."""
    rad_s = np.linalg.norm(S.half, axis=1)
    rad_t = np.linalg.norm(T.half, axis=1)
    useful_s = (S.first_child < 0) & (S.npts > 0)
    useful_t = (T.first_child < 0) & (T.npts > 0)

    def far(s, t):
        md = (2 * max(rad_s[s], rad_t[t]) + min(rad_s[s], rad_t[t])) * mac_scale
        d = S.center[s] - T.center[t]
        return md * md < d[0] * d[0] + d[1] * d[1] + d[2] * d[2]

    def fate(t, s):
        if far(s, t):
            return "far"
        if useful_t[t]:
            fc = S.first_child[s]
            if fc < 0 or not any(far(fc + o, t) for o in range(8)):
                return "close"
            return "desc_s"
        if (rad_t[t] < rad_s[s] or useful_t[t]) and not useful_s[s]:
            return "desc_s"
        return "desc_t"

    memo = {}

    def reached(t, s):
        if t == 0 and s == 0:
            return True
        if (t, s) in memo:
            return memo[(t, s)]
        ok = False
        if s != 0 and fate(t, S.parent[s]) == "desc_s" and reached(t, S.parent[s]):
            ok = True
        elif t != 0 and fate(T.parent[t], s) == "desc_t" and reached(T.parent[t], s):
            ok = True
        memo[(t, s)] = ok
        return ok

    bad = [("far", t, s) for t, s in zip(ft, fs) if fate(t, s) != "far" or not reached(t, s)]
    bad += [("close", t, s) for t, s in zip(ct, cs) if fate(t, s) != "close" or not reached(t, s)]
    return bad


@pytest.mark.parametrize("shared", [False, True])
def test_interaction_lists_follow_the_descent_rule(orc, shared):
    """In the shared frame source and target boxes of one level tie (R_t == R_s), the case where
    the rule's strict '<' decides which side descends; the planted '<=' mistake must change the
    lists and be caught."""
    rng = np.random.default_rng(14)
    src = shell_sources(rng, 2500)
    tgt = rng.uniform(-2, 2, (800, 3))
    S = orc.source_tree(src, 12)
    T = orc.target_tree(tgt, 10, root=(np.zeros(3), np.full(3, 3.0)) if shared else None)
    (ft, fs), (ct, cs) = orc.interaction_lists(S, T)
    assert len(ft) > 1000 and len(ct) > 100
    assert _rule_violations(S, T, ft, fs, ct, cs) == []
    if shared:
        rad_s = np.linalg.norm(S.half, axis=1)
        rad_t = np.linalg.norm(T.half, axis=1)
        assert np.sum(rad_s[fs] == rad_t[ft]) > 100        # same-size boxes tie in this frame
        (wt, ws), (wct, wcs) = orc.interaction_lists(S, T, planted_tie_mistake=True)
        assert set(zip(wt, ws)) != set(zip(ft, fs))
        assert _rule_violations(S, T, wt, ws, wct, wcs) != []

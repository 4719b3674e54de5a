"""ctypes front end of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT: only tests/, __graft_entry__.smoke()
and the cpu_baseline leg of bench.py import this module, as the checker.  The oracle
restates the reference FMM (FMM/FMMGPU of iamwhoamianamy/bachelors, REAL_IS_DOUBLE);
see the header of fmm_oracle.cpp for what pins each part.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

# flag bits (fmm_oracle.cpp: enum Flags)
SECTORAL_FIX = 1
FACT_FIX = 2
ANALYTIC_CURL = 4
P2P = 8
M2L_FULL = 16
FLOAT_EPS = 32
FAITHFUL = 0
CORRECTED = SECTORAL_FIX | FACT_FIX | ANALYTIC_CURL | P2P | M2L_FULL

OP_M2M, OP_M2L, OP_L2L = 0, 1, 2

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_sp = ctypes.POINTER(ctypes.c_size_t)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_tree_build.restype = ctypes.c_void_p
        L.orc_lists_build.restype = ctypes.c_void_p
        L.orc_fmm_run.restype = ctypes.c_void_p
        L.orc_fmm_run_bem.restype = ctypes.c_void_p
        for name in ("orc_tree_free", "orc_tree_num_nodes", "orc_tree_num_slots", "orc_tree_get",
                     "orc_lists_free", "orc_lists_counts", "orc_lists_get", "orc_fmm_free", "orc_fmm_get_A",
                     "orc_direct_bem",
                     "orc_treecode",
                     "orc_fmm_num_slots", "orc_fmm_get", "orc_fmm_times", "orc_fmm_counts",
                     "orc_fmm_local", "orc_fmm_multipole"):
            getattr(L, name).argtypes = None
        _lib = L
    return _lib


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops (the results do not depend on it)."""
    lib().orc_set_threads(int(n))


def max_threads():
    return int(lib().orc_max_threads())


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def nlen(p):
    return (p + 1) * (p + 1)


def solid_harmonics(p, v, regular=True, flags=FAITHFUL):
    """Reference Legendre route (harmonics.cpp:20-56)."""
    out = np.zeros(nlen(p))
    v = np.ascontiguousarray(v, dtype=np.float64)
    lib().orc_solid_harmonics(p, _d(v), int(regular), flags, _d(out))
    return out


def solid_harmonics_cart(p, v, regular=True):
    """Independent Cartesian-recurrence route (corrected: includes m = p)."""
    out = np.zeros(nlen(p))
    v = np.ascontiguousarray(v, dtype=np.float64)
    lib().orc_solid_harmonics_cart(p, _d(v), int(regular), _d(out))
    return out


def regular_grad(p, v):
    out = np.zeros((nlen(p), 3))
    v = np.ascontiguousarray(v, dtype=np.float64)
    lib().orc_regular_grad(p, _d(v), _d(out))
    return out


def m2m_complex_matrix(p, t, flags=FAITHFUL):
    L = nlen(p)
    re = np.zeros((L, L))
    im = np.zeros((L, L))
    t = np.ascontiguousarray(t, dtype=np.float64)
    lib().orc_m2m_complex_matrix(p, _d(t), flags, _d(re), _d(im))
    return re + 1j * im


def operator(kind, p, t, flags=FAITHFUL):
    """Real-basis translation operator T, x_out = T @ x_in."""
    L = nlen(p)
    T = np.zeros((L, L))
    t = np.ascontiguousarray(t, dtype=np.float64)
    lib().orc_operator(kind, p, _d(t), flags, _d(T))
    return T


def m2m_matrix_operator(p, t, flags=FAITHFUL):
    L = nlen(p)
    T = np.zeros((L, L))
    t = np.ascontiguousarray(t, dtype=np.float64)
    lib().orc_m2m_matrix_operator(p, _d(t), flags, _d(T))
    return T


def translate(kind, p, t, x, flags=FAITHFUL):
    out = np.zeros(nlen(p))
    t = np.ascontiguousarray(t, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    lib().orc_translate(kind, p, _d(t), flags, _d(x), _d(out))
    return out


def bounding_box(P):
    P = np.ascontiguousarray(P, dtype=np.float64)
    c = np.zeros(3)
    h = np.zeros(3)
    lib().orc_bounding_box(ctypes.c_size_t(len(P)), _d(P), _d(c), _d(h))
    return c, h


class Tree:
    """Reference octree (insert/subdivide restated) in canonical BFS numbering."""

    def __init__(self, P, thr, root=None, max_depth=21):
        self.P = np.ascontiguousarray(P, dtype=np.float64)
        if root is None:
            mode, c, h = 0, np.zeros(3), np.zeros(3)
        else:
            mode = 1
            c = np.ascontiguousarray(root[0], dtype=np.float64)
            h = np.ascontiguousarray(root[1], dtype=np.float64)
        L = lib()
        self.h = ctypes.c_void_p(L.orc_tree_build(ctypes.c_size_t(len(self.P)), _d(self.P), mode, _d(c), _d(h),
                                                  thr, max_depth))
        n = L.orc_tree_num_nodes(self.h)
        ns = L.orc_tree_num_slots(self.h)
        self.level = np.zeros(n, np.int32)
        self.parent = np.zeros(n, np.int32)
        self.first_child = np.zeros(n, np.int32)
        self.begin = np.zeros(n, np.int32)
        self.end = np.zeros(n, np.int32)
        self.npts = np.zeros(n, np.int32)
        self.center = np.zeros((n, 3))
        self.half = np.zeros((n, 3))
        self.slots = np.zeros(ns, np.int32)
        L.orc_tree_get(self.h, _i(self.level), _i(self.parent), _i(self.first_child), _i(self.begin),
                       _i(self.end), _i(self.npts), _d(self.center), _d(self.half), _i(self.slots))

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.orc_tree_free(self.h)
            self.h = None


def source_tree(P, cap, max_depth=21):
    return Tree(P, cap, root=(np.zeros(3), np.full(3, 3.0)), max_depth=max_depth)


def target_tree(P, cap, root=None, max_depth=21):
    return Tree(P, cap + 1, root=root, max_depth=max_depth)


def interaction_lists(S, T, mac_scale=1.0, planted_tie_mistake=False):
    """Far and close pairs (target, source) in BFS ids.  planted_tie_mistake: the rule with "<="
    in place of "<" (a deliberate error, for showing that the list checks catch it)."""
    L = lib()
    build = L.orc_lists_build_tie_s if planted_tie_mistake else L.orc_lists_build
    build.restype = ctypes.c_void_p
    h = ctypes.c_void_p(build(S.h, T.h, ctypes.c_double(mac_scale)))
    nf = ctypes.c_size_t()
    nc = ctypes.c_size_t()
    L.orc_lists_counts(h, ctypes.byref(nf), ctypes.byref(nc))
    ft = np.zeros(nf.value, np.int32)
    fs = np.zeros(nf.value, np.int32)
    ct = np.zeros(nc.value, np.int32)
    cs = np.zeros(nc.value, np.int32)
    L.orc_lists_get(h, S.h, T.h, _i(ft), _i(fs), _i(ct), _i(cs))
    L.orc_lists_free(h)
    return (ft, fs), (ct, cs)


class _Cfg(ctypes.Structure):
    _fields_ = [("p", ctypes.c_int), ("cap_s", ctypes.c_int), ("cap_t", ctypes.c_int), ("flags", ctypes.c_int),
                ("target_root_mode", ctypes.c_int), ("max_depth", ctypes.c_int),
                ("target_root_c", ctypes.c_double * 3), ("target_root_h", ctypes.c_double * 3),
                ("mac_scale", ctypes.c_double), ("current", ctypes.c_double), ("fd_eps", ctypes.c_double),
                ("problem", ctypes.c_int), ("source_root_c", ctypes.c_double * 3),
                ("source_root_h", ctypes.c_double * 3)]


class FMMRun:
    """One reference FMM evaluation (FastMultipoleSolver restated)."""

    def __init__(self, src, w, tgt, p=10, cap_s=32, cap_t=32, flags=FAITHFUL, J=None, target_root=None,
                 mac_scale=1.0, current=1.0, max_depth=21, fd_eps=1e-3, bem=None, source_root=None):
        """bem = (B, normal), each (N, 3): Problem::BEM with BEMQuadrature sources; the source root
        is then Box((0,0,0), (1.1, 1.1, 1.6)) (multipole_solver.cpp:668-669) unless given."""
        src = np.ascontiguousarray(src, dtype=np.float64)
        tgt = np.ascontiguousarray(tgt, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        c = _Cfg()
        c.p, c.cap_s, c.cap_t, c.flags = p, cap_s, cap_t, flags
        c.max_depth = max_depth
        if target_root is None:
            c.target_root_mode = 0
        else:
            c.target_root_mode = 1
            for a in range(3):
                c.target_root_c[a] = target_root[0][a]
                c.target_root_h[a] = target_root[1][a]
        c.mac_scale = mac_scale
        c.current = current
        c.fd_eps = fd_eps
        c.problem = 0 if bem is None else 1
        if source_root is None:
            source_root = (np.zeros(3), np.array([1.1, 1.1, 1.6]) if bem is not None else np.full(3, 3.0))
        for a in range(3):
            c.source_root_c[a] = source_root[0][a]
            c.source_root_h[a] = source_root[1][a]
        Jp = None
        if J is not None:
            J = np.ascontiguousarray(J, dtype=np.float64)
            Jp = _d(J)
        L = lib()
        self.p = p
        if bem is None:
            self.h = ctypes.c_void_p(L.orc_fmm_run(ctypes.byref(c), ctypes.c_size_t(len(src)), _d(src), _d(w), Jp,
                                                   ctypes.c_size_t(len(tgt)), _d(tgt)))
        else:
            Bv = np.ascontiguousarray(bem[0], dtype=np.float64)
            nv = np.ascontiguousarray(bem[1], dtype=np.float64)
            self._bem = (Bv, nv)
            self.h = ctypes.c_void_p(L.orc_fmm_run_bem(ctypes.byref(c), ctypes.c_size_t(len(src)), _d(src), _d(w),
                                                       _d(Bv), _d(nv), ctypes.c_size_t(len(tgt)), _d(tgt)))
        n = L.orc_fmm_num_slots(self.h)
        self.slots = np.zeros(n, np.int32)
        self.B = np.zeros((n, 3))
        self.Bfar = np.zeros((n, 3))
        self.Bnear = np.zeros((n, 3))
        L.orc_fmm_get(self.h, _i(self.slots), _d(self.B), _d(self.Bfar), _d(self.Bnear))
        self.A = np.zeros((n, 3))
        self.Afar = np.zeros((n, 3))
        self.Anear = np.zeros((n, 3))
        L.orc_fmm_get_A(self.h, _d(self.A), _d(self.Afar), _d(self.Anear))
        t = np.zeros(7)
        L.orc_fmm_times(self.h, _d(t))
        self.times = dict(zip(["tree", "p2m", "m2m", "lists", "m2l", "l2l", "eval"], t.tolist()))
        cnt = np.zeros(4, np.uint64)
        L.orc_fmm_counts(self.h, cnt.ctypes.data_as(_sp))
        self.counts = dict(zip(["source_nodes", "target_nodes", "far_pairs", "close_pairs"], cnt.tolist()))

    def local(self, bfs_id):
        out = np.zeros(3 * nlen(self.p))
        lib().orc_fmm_local(self.h, int(bfs_id), _d(out))
        return out.reshape(3, -1)

    def multipole(self, bfs_id):
        out = np.zeros(3 * nlen(self.p))
        lib().orc_fmm_multipole(self.h, int(bfs_id), _d(out))
        return out.reshape(3, -1)

    def B_input_order(self, nt):
        out = np.full((nt, 3), np.nan)
        out[self.slots] = self.B
        return out

    def A_input_order(self, nt):
        out = np.full((nt, 3), np.nan)
        out[self.slots] = self.A
        return out

    def treecode(self, x, node=0):
        """MultipoleSolver::calcB / calcA (current, point) at each row of x: the treecode M2P
        recursion from source node `node` (BFS id; 0 = root), FD curl for B."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        B = np.zeros((len(x), 3))
        A = np.zeros((len(x), 3))
        lib().orc_treecode(self.h, int(node), ctypes.c_size_t(len(x)), _d(x), _d(B), _d(A))
        return B, A

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.orc_fmm_free(self.h)
            self.h = None


def direct_B(src, w, tgt, J=None, current=1.0):
    src = np.ascontiguousarray(src, dtype=np.float64)
    tgt = np.ascontiguousarray(tgt, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    B = np.zeros((len(tgt), 3))
    Jp = None
    if J is not None:
        J = np.ascontiguousarray(J, dtype=np.float64)
        Jp = _d(J)
    lib().orc_direct_B(ctypes.c_size_t(len(src)), _d(src), _d(w), Jp, ctypes.c_size_t(len(tgt)), _d(tgt),
                       ctypes.c_double(current), _d(B))
    return B


def direct_A(src, w, tgt, J=None, current=1.0):
    """A(x) = current / (4 pi) sum w J / |x - y| (the reference's calcA scaling, no MU0)."""
    src = np.ascontiguousarray(src, dtype=np.float64)
    tgt = np.ascontiguousarray(tgt, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    A = np.zeros((len(tgt), 3))
    Jp = None
    if J is not None:
        J = np.ascontiguousarray(J, dtype=np.float64)
        Jp = _d(J)
    lib().orc_direct_A(ctypes.c_size_t(len(src)), _d(src), _d(w), Jp, ctypes.c_size_t(len(tgt)), _d(tgt),
                       ctypes.c_double(current), _d(A))
    return A


def direct_bem(src, w, Bv, nrm, tgt):
    """math::calcBEMIntegral at each target (integration.cpp:191-215)."""
    src = np.ascontiguousarray(src, dtype=np.float64)
    tgt = np.ascontiguousarray(tgt, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    Bv = np.ascontiguousarray(Bv, dtype=np.float64)
    nrm = np.ascontiguousarray(nrm, dtype=np.float64)
    out = np.zeros((len(tgt), 3))
    lib().orc_direct_bem(ctypes.c_size_t(len(src)), _d(src), _d(w), _d(Bv), _d(nrm), ctypes.c_size_t(len(tgt)),
                         _d(tgt), _d(out))
    return out


def direct_grad_w(src, w, tgt, Bstar, J=None, current=1.0):
    src = np.ascontiguousarray(src, dtype=np.float64)
    tgt = np.ascontiguousarray(tgt, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    Bstar = np.ascontiguousarray(Bstar, dtype=np.float64)
    g = np.zeros(len(src))
    Jp = None
    if J is not None:
        J = np.ascontiguousarray(J, dtype=np.float64)
        Jp = _d(J)
    lib().orc_direct_grad_w(ctypes.c_size_t(len(src)), _d(src), _d(w), Jp, ctypes.c_size_t(len(tgt)), _d(tgt),
                            ctypes.c_double(current), _d(Bstar), _d(g))
    return g

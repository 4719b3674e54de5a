"""MI355X-native Biot-Savart fast multipole method (drop-in for the reference
FastMultipoleSolver of iamwhoamianamy/bachelors, FMM/FMMGPU).

The compute runs in the HIP library libbachelors_fmm.so (built in-tree for gfx950 by
bachelors_amd.build); this module is a thin ctypes front end over its C ABI
(include/fmm_capi.h).  There is no CPU fallback: importing works anywhere, but every
computing call raises if the library or a GPU is missing.
"""
import ctypes
import enum
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbachelors_fmm.so")

FMM_OK = 0
PREC_FP64, PREC_FP32, PREC_BF16X3, PREC_F16X3 = 0, 1, 2, 3
FRAME_REFERENCE, FRAME_SHARED = 0, 1
NEAR_NONE, NEAR_P2P, NEAR_TREECODE = 0, 1, 2
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3, "f16x3": PREC_F16X3}


class M2MAlg(enum.IntEnum):          # multipole_solver_enums.hpp:3-10
    NoTranslation = 0
    ComplexTranslation = 1
    RealTranslation = 2
    Layers = 3
    Matrices = 4


class Device(enum.IntEnum):          # multipole_solver_enums.hpp:12-17
    CPU = 0
    GPU = 1
    Adaptive = 2


class M2LAlg(enum.IntEnum):          # multipole_solver_enums.hpp:20-24
    ComplexTranslation = 0
    Matrices = 1


class Problem(enum.IntEnum):         # multipole_solver_enums.hpp:26-30
    BioSavartLaplace = 0
    BEM = 1


class FMMError(RuntimeError):
    pass


class Config(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int), ("source_leaf_capacity", ctypes.c_int),
                ("target_leaf_capacity", ctypes.c_int), ("precision", ctypes.c_int), ("frame", ctypes.c_int),
                ("m2l_full", ctypes.c_int), ("max_depth", ctypes.c_int), ("near_field", ctypes.c_int),
                ("mac_scale", ctypes.c_double), ("source_root_center", ctypes.c_double * 3),
                ("source_root_half", ctypes.c_double * 3)]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("source_nodes", "target_nodes", "source_levels", "target_levels",
                                            "source_slots", "target_slots", "dropped_sources", "dropped_targets")] + \
               [(n, ctypes.c_longlong) for n in ("far_pairs", "close_pairs", "m2l_classes", "near_interactions")] + \
               [(n, ctypes.c_int) for n in ("order", "padded_length", "bfs_steps")] + \
               [(n, ctypes.c_double) for n in ("t_tree", "t_p2m", "t_m2m", "t_lists", "t_tables", "t_m2l", "t_l2l",
                                               "t_l2p", "t_p2p", "t_total", "t_m2l_gemm")] + \
               [(n, ctypes.c_int) for n in ("source_tree_sorted", "target_tree_sorted", "source_leaves",
                                            "target_leaves")] + \
               [(n, ctypes.c_double) for n in ("t_to_m2l_gemm", "t_after_m2l_gemm", "t_lists_far")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _share_torch_hip_runtime():
    """torch ships its own copy of the HIP runtime under the same soname (libamdhip64.so.7) and
    loads it by file name.  Loaded first, that copy also serves this library (the dynamic loader
    resolves the soname to it); loaded after the system copy, it is a second runtime in the
    process, and torch then reports no HIP device.  So the torch copy, where torch is installed,
    is loaded before this library, without importing torch."""
    import importlib.util
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return
    for d in (spec.submodule_search_locations or []) if spec else []:
        path = os.path.join(d, "lib", "libamdhip64.so")
        if os.path.exists(path):
            try:
                ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
            except OSError:
                pass   # unusable copy: the library then loads the system runtime
            return


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FMMError("HIP library missing: %s (run python -m bachelors_amd.build)" % LIB_PATH)
        _share_torch_hip_runtime()
        L = ctypes.CDLL(LIB_PATH)
        L.fmm_last_error.restype = ctypes.c_char_p
        L.fmm_last_error.argtypes = [_vp]
        L.fmm_create.argtypes = [ctypes.POINTER(_vp), ctypes.POINTER(Config)]
        L.fmm_destroy.argtypes = [_vp]
        L.fmm_config_init.argtypes = [ctypes.POINTER(Config)]
        for name in ("fmm_build_trees", "fmm_leaf_expansions", "fmm_upward", "fmm_downward", "fmm_synchronize"):
            getattr(L, name).argtypes = [_vp]
        L.fmm_upward_with.argtypes = [_vp, ctypes.c_int]
        L.fmm_set_sources.argtypes = [_vp, ctypes.c_size_t, _vp, _vp, _vp]
        L.fmm_set_sources_dev.argtypes = [_vp, ctypes.c_size_t, _vp, _vp, _vp]
        L.fmm_set_targets.argtypes = [_vp, ctypes.c_size_t, _vp]
        L.fmm_set_targets_dev.argtypes = [_vp, ctypes.c_size_t, _vp]
        L.fmm_set_weights_dev.argtypes = [_vp, _vp]
        L.fmm_set_bem_sources.argtypes = [_vp, ctypes.c_size_t, _vp, _vp, _vp, _vp]
        L.fmm_set_bem_sources_dev.argtypes = [_vp, ctypes.c_size_t, _vp, _vp, _vp, _vp]
        L.fmm_evaluate_bem.argtypes = [_vp, _vp]
        L.fmm_evaluate_bem_dev.argtypes = [_vp, _vp]
        L.fmm_set_stream.argtypes = [_vp, _vp]
        L.fmm_evaluate_B.argtypes = [_vp, ctypes.c_double, _vp]
        L.fmm_evaluate_B_dev.argtypes = [_vp, ctypes.c_double, _vp]
        L.fmm_run_B_dev.argtypes = [_vp, ctypes.c_double, _vp]
        L.fmm_leaf_order.argtypes = [_vp, _vp]
        L.fmm_adjoint_weights_dev.argtypes = [_vp, ctypes.c_double, _vp, _vp]
        L.fmm_direct_B_dev.argtypes = [_vp, ctypes.c_double, ctypes.c_size_t, _vp, _vp]
        L.fmm_direct_A_dev.argtypes = [_vp, ctypes.c_double, ctypes.c_size_t, _vp, _vp]
        L.fmm_evaluate_A.argtypes = [_vp, ctypes.c_double, _vp]
        L.fmm_evaluate_A_dev.argtypes = [_vp, ctypes.c_double, _vp]
        L.fmm_treecode.argtypes = [_vp, ctypes.c_double, ctypes.c_size_t, _vp, _vp, _vp]
        L.fmm_treecode_dev.argtypes = [_vp, ctypes.c_double, ctypes.c_size_t, _vp, _vp, _vp]
        L.fmm_get_stats.argtypes = [_vp, ctypes.POINTER(Stats)]
        L.fmm_tree_size.argtypes = [_vp, ctypes.c_int, _ip, _ip]
        L.fmm_tree_export.argtypes = [_vp, ctypes.c_int] + [_vp] * 8
        L.fmm_lists_size.argtypes = [_vp] + [ctypes.POINTER(ctypes.c_longlong)] * 3
        L.fmm_lists_export.argtypes = [_vp] + [_vp] * 4
        L.fmm_export_B_parts.argtypes = [_vp, _vp, _vp]
        L.fmm_export_expansion.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp]
        L.fmm_make_operator.argtypes = [ctypes.c_int, ctypes.c_int, _vp, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, _vp]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def default_config(**kw):
    c = Config()
    lib().fmm_config_init(ctypes.byref(c))
    for k, v in kw.items():
        if k in ("source_root_center", "source_root_half"):
            for a in range(3):
                getattr(c, k)[a] = v[a]
        elif k == "precision" and isinstance(v, str):
            c.precision = PRECISIONS[v]
        else:
            setattr(c, k, v)
    return c


class FMM:
    """One FMM context (fmm_ctx): owns device memory, a HIP stream and the trees."""

    def __init__(self, **cfg):
        self.cfg = default_config(**cfg)
        h = _vp()
        rc = lib().fmm_create(ctypes.byref(h), ctypes.byref(self.cfg))
        if rc != FMM_OK:
            raise FMMError("fmm_create failed (code %d): check the configuration and the GPU" % rc)
        self.h = h
        self.nt = 0
        self.ns = 0

    def _check(self, rc):
        if rc != FMM_OK:
            raise FMMError("code %d: %s" % (rc, lib().fmm_last_error(self.h).decode()))

    def close(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.fmm_destroy(self.h)
            self.h = None

    __del__ = close

    # ---- inputs (host numpy)
    def set_sources(self, xyz, w, J=None):
        xyz = np.ascontiguousarray(xyz, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        J = None if J is None else np.ascontiguousarray(J, np.float64)
        self._keep = (xyz, w, J)
        self.ns = len(xyz)
        self._check(lib().fmm_set_sources(self.h, len(xyz), _ptr(xyz), _ptr(w), _ptr(J)))

    def set_bem_sources(self, xyz, w, B, normal):
        """Problem::BEM sources (BEMQuadrature): weight, field B and normal per point."""
        xyz, w, B, normal = (np.ascontiguousarray(a, np.float64) for a in (xyz, w, B, normal))
        self._keep = (xyz, w, B, normal)
        self.ns = len(xyz)
        self._check(lib().fmm_set_bem_sources(self.h, len(xyz), _ptr(xyz), _ptr(w), _ptr(B), _ptr(normal)))

    def evaluate_bem(self):
        """calcBEM at every target, input order."""
        out = np.zeros((self.nt, 3))
        self._check(lib().fmm_evaluate_bem(self.h, _ptr(out)))
        return out

    def set_targets(self, xyz):
        xyz = np.ascontiguousarray(xyz, np.float64)
        self.nt = len(xyz)
        self._check(lib().fmm_set_targets(self.h, len(xyz), _ptr(xyz)))

    # ---- inputs (device pointers, e.g. torch tensors' data_ptr())
    def set_sources_dev(self, n, xyz_ptr, w_ptr, J_ptr=None):
        self.ns = n
        self._check(lib().fmm_set_sources_dev(self.h, n, xyz_ptr, w_ptr, J_ptr))

    def set_targets_dev(self, m, xyz_ptr):
        self.nt = m
        self._check(lib().fmm_set_targets_dev(self.h, m, xyz_ptr))

    def set_weights_dev(self, w_ptr):
        self._check(lib().fmm_set_weights_dev(self.h, w_ptr))

    def set_stream(self, stream_ptr):
        self._check(lib().fmm_set_stream(self.h, stream_ptr))

    # ---- stages
    def build_trees(self):
        self._check(lib().fmm_build_trees(self.h))

    def leaf_expansions(self):
        self._check(lib().fmm_leaf_expansions(self.h))

    def upward(self, algorithm=None):
        """M2M; algorithm: an M2MAlg (default Matrices, the MFMA GEMM path)."""
        if algorithm is None:
            self._check(lib().fmm_upward(self.h))
        else:
            self._check(lib().fmm_upward_with(self.h, int(algorithm)))

    def downward(self):
        self._check(lib().fmm_downward(self.h))

    def evaluate_B(self, current=1.0):
        B = np.zeros((self.nt, 3))
        self._check(lib().fmm_evaluate_B(self.h, current, _ptr(B)))
        return B

    def evaluate_B_dev(self, current, B_ptr):
        self._check(lib().fmm_evaluate_B_dev(self.h, current, B_ptr))

    def evaluate_A(self, current=1.0):
        """Vector potential at every target, input order (calcA(current) scaling: no MU0)."""
        A = np.zeros((self.nt, 3))
        self._check(lib().fmm_evaluate_A(self.h, current, _ptr(A)))
        return A

    def evaluate_A_dev(self, current, A_ptr):
        self._check(lib().fmm_evaluate_A_dev(self.h, current, A_ptr))

    def treecode(self, x, current=1.0):
        """Single-point treecode B and A (MultipoleSolver::calcB / calcA(current, point)) at each row."""
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        B = np.zeros((len(x), 3))
        A = np.zeros((len(x), 3))
        self._check(lib().fmm_treecode(self.h, current, len(x), _ptr(x), _ptr(B), _ptr(A)))
        return B, A

    def treecode_dev(self, current, m, xyz_ptr, B_ptr, A_ptr):
        self._check(lib().fmm_treecode_dev(self.h, current, m, xyz_ptr, B_ptr, A_ptr))

    def run_B_dev(self, current, B_ptr):
        self._check(lib().fmm_run_B_dev(self.h, current, B_ptr))

    def run(self, current=1.0):
        self.build_trees()
        self.leaf_expansions()
        self.upward()
        self.downward()
        return self.evaluate_B(current)

    def synchronize(self):
        self._check(lib().fmm_synchronize(self.h))

    def adjoint_weights_dev(self, current, dLdB_ptr, dLdw_ptr):
        self._check(lib().fmm_adjoint_weights_dev(self.h, current, dLdB_ptr, dLdw_ptr))

    def direct_B_dev(self, current, m, xyz_ptr, B_ptr):
        self._check(lib().fmm_direct_B_dev(self.h, current, m, xyz_ptr, B_ptr))

    def direct_A_dev(self, current, m, xyz_ptr, A_ptr):
        self._check(lib().fmm_direct_A_dev(self.h, current, m, xyz_ptr, A_ptr))

    def leaf_order(self):
        n = self.tree(1)["slots"].shape[0]
        out = np.zeros(n, np.int64)
        self._check(lib().fmm_leaf_order(self.h, _ptr(out)))
        return out

    def stats(self):
        s = Stats()
        self._check(lib().fmm_get_stats(self.h, ctypes.byref(s)))
        return s.as_dict()

    # ---- parity export
    def tree(self, which):
        nn = ctypes.c_int()
        ns = ctypes.c_int()
        self._check(lib().fmm_tree_size(self.h, which, ctypes.byref(nn), ctypes.byref(ns)))
        n = nn.value
        d = {k: np.zeros(n, np.int32) for k in ("level", "parent", "first_child", "begin", "end", "npts")}
        d["center"] = np.zeros((n, 3))
        d["slots"] = np.zeros(ns.value, np.int32)
        self._check(lib().fmm_tree_export(self.h, which, _ptr(d["level"]), _ptr(d["parent"]),
                                          _ptr(d["first_child"]), _ptr(d["begin"]), _ptr(d["end"]),
                                          _ptr(d["npts"]), _ptr(d["center"]), _ptr(d["slots"])))
        return d

    def lists(self):
        nf = ctypes.c_longlong()
        nc = ctypes.c_longlong()
        ncl = ctypes.c_longlong()
        self._check(lib().fmm_lists_size(self.h, ctypes.byref(nf), ctypes.byref(nc), ctypes.byref(ncl)))
        ft = np.zeros(nf.value, np.int32)
        fs = np.zeros(nf.value, np.int32)
        ct = np.zeros(nc.value, np.int32)
        cs = np.zeros(nc.value, np.int32)
        self._check(lib().fmm_lists_export(self.h, _ptr(ft), _ptr(fs), _ptr(ct), _ptr(cs)))
        return (ft, fs), (ct, cs), ncl.value

    def B_parts(self):
        n = self.tree(1)["slots"].shape[0]
        Bf = np.zeros((n, 3))
        Bn = np.zeros((n, 3))
        self._check(lib().fmm_export_B_parts(self.h, _ptr(Bf), _ptr(Bn)))
        return Bf, Bn

    def expansion(self, which, node):
        p = self.cfg.order
        out = np.zeros(3 * (p + 1) ** 2)
        self._check(lib().fmm_export_expansion(self.h, which, int(node), _ptr(out)))
        return out.reshape(3, -1)


def make_operator(kind, order, t, sig_in=1.0, sig_out=1.0, m2l_full=1):
    """Real-basis translation operator built by the GPU table generator (row = output)."""
    L = (order + 1) ** 2
    out = np.zeros((L, L))
    t = np.ascontiguousarray(t, np.float64)
    rc = lib().fmm_make_operator(kind, order, _ptr(t), sig_in, sig_out, m2l_full, _ptr(out))
    if rc != FMM_OK:
        raise FMMError("fmm_make_operator failed (code %d)" % rc)
    return out


class FastMultipoleSolver:
    """Mirror of the reference FastMultipoleSolver (fast_multipole_solver.hpp:13-98).

    quadratures: (N, 4) array of (x, y, z, weight) -- the reference Quadrature (position
    plus weight = volume * basis weight, quadrature.cpp:7-10) -- or an (N, 3) position
    array together with `weights`.  points: (M, 3) targets.  The current direction of a
    quadrature is phi-hat (integration.cpp:155) unless `currents` (N, 3) is given.
    """

    def __init__(self, quadratures, points, problem=Problem.BioSavartLaplace, quadratureOctreeLeafCapacity=1000,
                 calculationPointOctreeLeafCapacity=100, weights=None, currents=None, order=10,
                 precision="fp64", frame=FRAME_REFERENCE, mac_scale=1.0, m2l_full=True, near_field=NEAR_P2P,
                 max_depth=None, bem_B=None, bem_normal=None):
        """Problem.BEM takes BEMQuadrature sources: quadratures (N, 4) or positions + weights, with
        bem_B and bem_normal (N, 3); their tree is rooted at Box((0,0,0), (1.1, 1.1, 1.6))
        (multipole_solver.cpp:668-669)."""
        problem = Problem(problem)
        if problem == Problem.BEM and (bem_B is None or bem_normal is None):
            raise FMMError("Problem.BEM needs bem_B and bem_normal for every quadrature")
        self.problem = problem
        q = np.asarray(quadratures, np.float64)
        if weights is None:
            xyz, w = q[:, :3], q[:, 3]
        else:
            xyz, w = q[:, :3], np.asarray(weights, np.float64)
        self.harmonicOrder = order
        self.harmonicLength = (order + 1) ** 2
        self.quadratureOctreeLeafCapacity = quadratureOctreeLeafCapacity
        self.calculationPointOctreeLeafCapacity = calculationPointOctreeLeafCapacity
        self.log = False
        self._points = np.ascontiguousarray(points, np.float64)
        extra = {} if max_depth is None else {"max_depth": max_depth}
        if problem == Problem.BEM:
            extra.update(source_root_center=(0.0, 0.0, 0.0), source_root_half=(1.1, 1.1, 1.6))
        self._fmm = FMM(order=order, source_leaf_capacity=quadratureOctreeLeafCapacity,
                        target_leaf_capacity=calculationPointOctreeLeafCapacity, precision=precision, frame=frame,
                        mac_scale=mac_scale, m2l_full=int(m2l_full), near_field=int(near_field), **extra)
        if problem == Problem.BEM:
            self._fmm.set_bem_sources(xyz, w, bem_B, bem_normal)
        else:
            self._fmm.set_sources(xyz, w, currents)
        self._fmm.set_targets(self._points)
        self._fmm.build_trees()              # the constructor builds both trees (initTrees)
        self._leaves_ready = False
        self._multipoles_ready = False
        self._locals_ready = False

    def calcMultipoleExpansionsAtLeaves(self):
        self._fmm.leaf_expansions()
        self._leaves_ready = True

    def calclMultipoleExpansions(self, algorithm=M2MAlg.Matrices, device=Device.CPU):
        """M2M by the reference's algorithm choice (multipole_solver.cpp:102-139), all on the GPU:
        Matrices = MFMA GEMMs of per-(level, orientation) operators; ComplexTranslation,
        RealTranslation, Layers = the per-node translation formula; NoTranslation = P2M at every
        node.  The device argument is accepted for source compatibility."""
        if not self._leaves_ready:
            self.calcMultipoleExpansionsAtLeaves()
        self._fmm.upward(M2MAlg(algorithm))
        self._multipoles_ready = True

    def calcLocalMultipoleExpansions(self, algorithm=M2LAlg.ComplexTranslation, device=Device.CPU):
        if not self._multipoles_ready:
            raise FMMError("Multipoles are not ready!")      # fast_multipole_solver.cpp:100
        self._fmm.downward()
        self._locals_ready = True

    def getQuadratureOctreeNodeCount(self):
        """All source-tree nodes, empty children included (multipole_solver.cpp:97-100)."""
        return self._fmm.tree(0)["level"].shape[0]

    def calcB(self, current, point=None):
        """calcB(current): list of (point, B) in target-leaf order (FMMResult order,
        fast_multipole_solver.cpp:214-231).  calcB(current, point): the treecode B at one point
        (multipole_solver.cpp:694-700), after calclMultipoleExpansions."""
        if point is not None:
            self._need_multipoles()
            return self._fmm.treecode(np.asarray(point, np.float64), current)[0][0]
        B = self._fmm.evaluate_B(current)
        order = self._fmm.leaf_order()
        return [(self._points[i], B[i]) for i in order]

    def calcA(self, current, point=None):
        """calcA(current): list of (point, A) in target-leaf order (fast_multipole_solver.cpp:195-212,
        A = current / (4 pi) sum w J / |x - y|).  calcA(current, point): the treecode A at one point
        (multipole_solver.cpp:686-692)."""
        if point is not None:
            self._need_multipoles()
            return self._fmm.treecode(np.asarray(point, np.float64), current)[1][0]
        A = self._fmm.evaluate_A(current)
        order = self._fmm.leaf_order()
        return [(self._points[i], A[i]) for i in order]

    def calcBEM(self, current=1.0):
        """calcBEM (fast_multipole_solver.cpp:233-250): list of (point, value) in target-leaf order;
        as in the reference, current does not enter."""
        V = self._fmm.evaluate_bem()
        order = self._fmm.leaf_order()
        return [(self._points[i], V[i]) for i in order]

    def calcBEM_array(self, current=1.0):
        return self._fmm.evaluate_bem()

    def calcA_array(self, current):
        """A at every target in input order, (M, 3)."""
        return self._fmm.evaluate_A(current)

    def treecode(self, current, points):
        """calcB and calcA (current, point) at many points at once: (B, A), each (n, 3)."""
        self._need_multipoles()
        return self._fmm.treecode(points, current)

    def _need_multipoles(self):
        if not self._multipoles_ready:
            raise FMMError("Multipoles are not ready!")      # multipole_solver.cpp:689,697

    def calcB_array(self, current):
        """B at every target in input order, (M, 3)."""
        return self._fmm.evaluate_B(current)

    def stats(self):
        return self._fmm.stats()

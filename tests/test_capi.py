"""CPU: the C-ABI library loads and exports every function include/*.h declares, and the
host-side argument checks behave without a GPU (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|void|const char\*)\s+(fmm_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    import bachelors_amd
    L = bachelors_amd.lib()
    names = _declared_functions()
    assert len(names) >= 25
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_config_defaults_match_reference():
    import bachelors_amd
    c = bachelors_amd.default_config()
    assert c.order == 10                     # multipole_solver.hpp:25
    assert c.source_leaf_capacity == 1000    # fast_multipole_solver.hpp:29
    assert c.target_leaf_capacity == 100     # fast_multipole_solver.hpp:30
    assert c.mac_scale == 1.0
    assert list(c.source_root_half) == [3.0, 3.0, 3.0]   # multipole_solver.cpp:662-663
    assert list(c.source_root_center) == [0.0, 0.0, 0.0]
    # the reference's target tree (bounding-box root, math.cpp:39-57) and leaf order
    assert c.frame == bachelors_amd.FRAME_REFERENCE
    assert c.max_depth == 21 and c.precision == bachelors_amd.PREC_FP64


def test_drop_in_defaults_to_the_reference_frame():
    import inspect
    import bachelors_amd
    sig = inspect.signature(bachelors_amd.FastMultipoleSolver.__init__)
    assert sig.parameters["frame"].default == bachelors_amd.FRAME_REFERENCE
    assert sig.parameters["quadratureOctreeLeafCapacity"].default == 1000   # fast_multipole_solver.hpp:29
    assert sig.parameters["calculationPointOctreeLeafCapacity"].default == 100


def test_create_rejects_bad_arguments():
    import bachelors_amd
    L = bachelors_amd.lib()
    h = ctypes.c_void_p()
    for kw in ({"order": 0}, {"order": 99}, {"source_leaf_capacity": 1}, {"precision": 7}, {"frame": 3},
               {"mac_scale": 0.5}):
        c = bachelors_amd.default_config(**kw)
        assert L.fmm_create(ctypes.byref(h), ctypes.byref(c)) == 1
    assert L.fmm_create(None, None) == 1
    # a null context is rejected, never dereferenced
    assert L.fmm_build_trees(None) == 1
    assert L.fmm_last_error(None) == b"null context"


def test_cpp_drop_in_header_compiles_and_links():
    """The C++ mirror of FastMultipoleSolver builds against the library with the host compiler."""
    import subprocess
    d = os.path.join(ROOT, "tests", "cpp")
    subprocess.run(["make", "-s", "-B", "-C", d], check=True)
    assert os.path.exists(os.path.join(d, "drop_in_example"))


def test_one_hip_runtime_with_torch():
    """The library loaded before torch and torch imported after it map one HIP runtime between
    them (torch's bundled copy), in a fresh process: two copies in one process leave torch
    without a device."""
    import subprocess
    import sys
    code = ("import sys, re; sys.path.insert(0, '.'); import bachelors_amd as fm; fm.lib(); import torch; "
            "print(sorted(set(re.findall(r'(/\\S*libamdhip64\\S*)', open('/proc/self/maps').read()))))")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    maps = eval(r.stdout.strip().splitlines()[-1])
    assert len(maps) == 1, maps

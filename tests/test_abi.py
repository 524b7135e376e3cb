"""C ABI checks that need no GPU: libvsdmo.so loads and exports every function
declared in include/vsdmo.h; the C++ drop-in header compiles against it."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vsdmo.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(vsdmo_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("vsdmo_mesh_create", "vsdmo_opt_create", "vsdmo_opt_iterate", "vsdmo_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from volumesurfacediscretemeshoptimization_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding binds exactly the declared set
    assert set(_lib.EXPORTS) == set(declared_functions())


def test_cpp_wrapper_compiles():
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = os.path.join(ROOT, "tests", "cpp", "dropin_example.cpp")
    out = "/tmp/vsdmo_dropin_example"
    r = subprocess.run([gxx, "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-o", out,
                        "-L", os.path.join(ROOT, "volumesurfacediscretemeshoptimization_amd"), "-lvsdmo",
                        "-Wl,-rpath," + os.path.join(ROOT, "volumesurfacediscretemeshoptimization_amd")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_defaults():
    from volumesurfacediscretemeshoptimization_amd import _lib
    lib = _lib.load()
    p = _lib.Params()
    lib.vsdmo_params_default(ctypes.byref(p))
    assert (p.nq, p.depth, p.shrink, p.extent_mode) == (8, 3, 0.0, -1)
    mp = _lib.MeshParams()
    lib.vsdmo_mesh_params_default(ctypes.byref(mp))
    assert mp.coloring == 0 and mp.feature_angle_deg == 20.0


def test_invalid_arguments_rejected_without_gpu_work():
    """Argument validation happens before any device call."""
    import numpy as np
    from volumesurfacediscretemeshoptimization_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    xyz = np.zeros((4, 3), np.float32)
    bad = np.array([[0, 1, 2, 7]], np.int32)       # index out of range
    st = lib.vsdmo_mesh_create(_lib.VSDMO_TET, xyz.ctypes.data_as(_lib._fp), 4, bad.ctypes.data_as(_lib._ip), 1,
                               None, None, ctypes.byref(h))
    assert st == _lib.VSDMO_INVALID_ARG and not h.value
    assert b"out of range" in lib.vsdmo_last_error()
    dup = np.array([[0, 1, 1, 2]], np.int32)
    st = lib.vsdmo_mesh_create(_lib.VSDMO_TET, xyz.ctypes.data_as(_lib._fp), 4, dup.ctypes.data_as(_lib._ip), 1,
                               None, None, ctypes.byref(h))
    assert st == _lib.VSDMO_INVALID_ARG
    assert lib.vsdmo_opt_create(None, 0, 0.5, 1, None, ctypes.byref(h)) == _lib.VSDMO_INVALID_ARG


def test_struct_layouts_match_the_header():
    """the ctypes mirrors of the ABI structs have the C layout: the size and
    every field offset, as a C compiler lays out include/vsdmo.h"""
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    from volumesurfacediscretemeshoptimization_amd import _lib
    pairs = {"vsdmo_mesh_params": _lib.MeshParams, "vsdmo_params": _lib.Params, "vsdmo_counts": _lib.Counts,
             "vsdmo_stats": _lib.Stats, "vsdmo_batch_params": _lib.BatchParams,
             "vsdmo_batch_stats": _lib.BatchStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "vsdmo.h"', "int main(void) {"]
    for cname, py in pairs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src, exe = "/tmp/vsdmo_layout.c", "/tmp/vsdmo_layout"
    open(src, "w").write("\n".join(lines) + "\n")
    r = subprocess.run([gcc, "-I", os.path.join(ROOT, "include"), src, "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in pairs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)

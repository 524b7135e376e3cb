"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: tests compare the HIP path against it.  It is
never the thing measured or shipped.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

OR_TRI, OR_QUAD, OR_TET, OR_HEX = 0, 1, 2, 3
OR_COLOR_GREEDY, OR_COLOR_JPL = 0, 1
OR_EXTENT_MAX, OR_EXTENT_MIN = 0, 1
OR_HEX_SHAPE9, OR_HEX_CORNER8 = 0, 1

_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)
_bp = C.POINTER(C.c_ubyte)


class OrMesh(C.Structure):
    _fields_ = [
        ("elem", C.c_int), ("npe", C.c_int), ("nV", C.c_int), ("nE", C.c_int), ("flat", C.c_int),
        ("xyz", _fp), ("cells", _ip), ("new_of_old", _ip), ("old_of_new", _ip),
        ("nSurf", C.c_int), ("nFeature", C.c_int), ("nFree", C.c_int), ("nInner", C.c_int),
        ("nColorsInner", C.c_int), ("nColorsFree", C.c_int), ("nColorsFeature", C.c_int),
        ("colOffInner", _ip), ("colOffFree", _ip), ("colOffFeature", _ip), ("colorOf", _ip),
        ("nFaces", C.c_int), ("fpe", C.c_int), ("faces", _ip),
        ("ringStride", C.c_int), ("ringRow", _ip), ("ringVal", _ip), ("maxRing", C.c_int),
        ("nHE", C.c_int), ("heRow", _ip), ("heTarget", _ip), ("heOpp", _ip), ("heFace", _ip), ("heNext", _ip),
        ("maxHE", C.c_int), ("isBoundary1D", _bp), ("isFeature", _bp),
    ]


class BuildParams(C.Structure):
    _fields_ = [("coloring", C.c_int), ("feature_angle_deg", C.c_float)]


class DmoParams(C.Structure):
    _fields_ = [("nq", C.c_int), ("depth", C.c_int), ("grid_scale", C.c_float), ("shrink", C.c_float),
                ("extent_mode", C.c_int), ("fixed_boundary", C.c_int), ("nthreads", C.c_int),
                ("max_per_class", C.c_int), ("hex_metric", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        _lib = _configure(C.CDLL(LIB_PATH))
    return _lib


@contextlib.contextmanager
def use_library(path: str):
    """Runs the block against another build of the oracle (the mutation tests'
    copies with a planted mistake); the default library is restored after."""
    global _lib
    saved = lib()
    _lib = _configure(C.CDLL(path))
    try:
        yield _lib
    finally:
        _lib = saved


def _configure(L):
    """Sets the argument and result types of every exported oracle function."""
    L.or_cbrt.restype = C.c_float; L.or_cbrt.argtypes = [C.c_float]
    L.or_pow2_scale.restype = C.c_float; L.or_pow2_scale.argtypes = [C.c_float]
    for n in ("or_q_tet", "or_q_hex", "or_q_tri2d"):
        getattr(L, n).restype = C.c_float; getattr(L, n).argtypes = [_fp]
    for n in ("or_q_tet64", "or_q_hex64", "or_q_tri2d64"):
        getattr(L, n).restype = C.c_double; getattr(L, n).argtypes = [_fp]
    L.or_q_hex_m.restype = C.c_float; L.or_q_hex_m.argtypes = [_fp, C.c_int]
    L.or_q_hex64_m.restype = C.c_double; L.or_q_hex64_m.argtypes = [_fp, C.c_int]
    L.or_ring_score_hex_m.restype = C.c_float
    L.or_ring_score_hex_m.argtypes = [C.c_int, _fp, _fp, C.c_float, _fp, C.c_int]
    for n in ("or_ring_score_tet", "or_ring_score_hex", "or_ring_score_tri2d"):
        getattr(L, n).restype = C.c_float
        getattr(L, n).argtypes = [C.c_int, _fp, _fp, C.c_float, _fp]
    L.or_surface_points.restype = None
    L.or_surface_points.argtypes = [C.POINTER(OrMesh), C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_float, _fp]
    L.or_fit_local_surface.restype = None
    L.or_fit_local_surface.argtypes = [_fp, C.c_int, _fp, _fp, _fp]
    L.or_lu_solve3.restype = None
    L.or_lu_solve3.argtypes = [_fp, _fp, _fp]
    L.or_mesh_build.restype = C.c_int
    L.or_mesh_build.argtypes = [C.c_int, _fp, C.c_int, _ip, C.c_int, C.POINTER(BuildParams), C.POINTER(C.POINTER(OrMesh))]
    L.or_mesh_free.argtypes = [C.POINTER(OrMesh)]
    L.or_boundary_faces.restype = C.c_int
    L.or_boundary_faces.argtypes = [C.c_int, C.c_int, _ip, _ip]
    L.or_color.restype = C.c_int
    L.or_color.argtypes = [C.c_int, C.c_int, _ip, _ip, _ip]
    L.or_adjacency_range.restype = C.c_int
    L.or_adjacency_range.argtypes = [C.c_int, C.c_int, _ip, C.c_int, C.c_int, C.POINTER(_ip), C.POINTER(_ip)]
    L.or_dmo_iteration.restype = C.c_longlong
    L.or_dmo_iteration.argtypes = [C.POINTER(OrMesh), C.POINTER(DmoParams)]
    L.or_dmo_vertex.restype = C.c_int
    L.or_dmo_vertex.argtypes = [C.POINTER(OrMesh), C.POINTER(DmoParams), C.c_int, _fp]
    L.or_min_quality.restype = C.c_float
    L.or_min_quality.argtypes = [C.POINTER(OrMesh)]
    L.or_quality64.argtypes = [C.POINTER(OrMesh), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.or_element_qualities.argtypes = [C.POINTER(OrMesh), _fp]
    L.or_min_quality_m.restype = C.c_float
    L.or_min_quality_m.argtypes = [C.POINTER(OrMesh), C.c_int]
    L.or_quality64_m.argtypes = [C.POINTER(OrMesh), C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.or_element_qualities_m.argtypes = [C.POINTER(OrMesh), C.c_int, _fp]
    L.or_surface_init.restype = C.c_void_p
    L.or_surface_init.argtypes = [C.POINTER(OrMesh)]
    L.or_surface_init_ring.restype = C.c_void_p
    L.or_surface_init_ring.argtypes = [C.POINTER(OrMesh), C.c_int]
    L.or_surface_free.argtypes = [C.c_void_p]
    L.or_dmo_iteration_surface.restype = C.c_longlong
    L.or_dmo_iteration_surface.argtypes = [C.POINTER(OrMesh), C.c_void_p, C.POINTER(DmoParams), C.c_float,
                                           C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.or_dmo_iteration_surface_mesh.restype = C.c_longlong
    L.or_dmo_iteration_surface_mesh.argtypes = [C.POINTER(OrMesh), C.c_void_p, C.POINTER(DmoParams),
                                                C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.or_surface_mesh_qualities.argtypes = [C.POINTER(OrMesh), _fp]
    L.or_surface_mesh_min_quality.restype = C.c_float
    L.or_surface_mesh_min_quality.argtypes = [C.POINTER(OrMesh)]
    L.or_laplace.argtypes = [C.POINTER(OrMesh), C.c_int]
    for n in ("or_q_tri3", "or_q_quad"):
        getattr(L, n).restype = C.c_float; getattr(L, n).argtypes = [_fp]
    for n in ("or_q_tri3n", "or_q_quadn"):
        getattr(L, n).restype = C.c_float; getattr(L, n).argtypes = [_fp, _fp]
    for n in ("or_q_tri364", "or_q_quad64"):
        getattr(L, n).restype = C.c_double; getattr(L, n).argtypes = [_fp]
    L.or_surface_get.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.or_surface_nsectors.restype = C.c_int
    L.or_surface_nsectors.argtypes = [C.c_void_p, C.c_int]
    L.or_search_ring.restype = C.c_int
    L.or_search_ring.argtypes = [C.c_int, C.c_int, _fp, _fp, C.POINTER(DmoParams), _fp]
    L.or_surface_get_sectors.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.or_surface_search_vertex.restype = C.c_int
    L.or_surface_search_vertex.argtypes = [C.POINTER(OrMesh), C.c_void_p, C.POINTER(DmoParams), C.c_int, C.c_int,
                                           _fp, _fp]
    L.or_surface_project.argtypes = [C.POINTER(OrMesh), C.c_void_p, C.c_int, C.c_float, C.c_float, _fp]
    L.or_surface_update_nn.argtypes = [C.POINTER(OrMesh), C.c_void_p]
    L.or_surface_error.restype = C.c_int
    L.or_surface_error.argtypes = [C.c_void_p]
    L.or_learned_window.restype = C.c_int
    L.or_learned_window.argtypes = [C.POINTER(OrMesh), C.POINTER(DmoParams), C.c_int, _fp, _fp, C.c_int, C.c_float, _fp]
    return L


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def cbrt(x: float) -> float:
    return lib().or_cbrt(float(x))


def q_elem(kind: str, pts, fp64: bool = False, normal=None, hex_metric: int = 0) -> float:
    """kind: tet, hex, tri (2D), tri3 (surface triangle), quad (surface quad);
    normal: signed surface variants (tri3 / quad); hex_metric: OR_HEX_SHAPE9
    or OR_HEX_CORNER8 for hexes"""
    a, p = _f(pts)
    if kind == "hex" and hex_metric:
        return (lib().or_q_hex64_m if fp64 else lib().or_q_hex_m)(p, int(hex_metric))
    name = {"tet": "or_q_tet", "hex": "or_q_hex", "tri": "or_q_tri2d", "tri3": "or_q_tri3", "quad": "or_q_quad"}[kind]
    if normal is not None:
        n, npp = _f(normal)
        return getattr(lib(), name + "n")(p, npp)
    return getattr(lib(), name + ("64" if fp64 else ""))(p)


def ring_score(kind: str, ring, x0, sigma, p, hex_metric: int = 0) -> float:
    r, rp = _f(ring)
    x, xp = _f(x0)
    q, qp = _f(p)
    per = {"tet": 9, "hex": 21, "tri": 6}[kind]
    n = r.size // per
    if kind == "hex" and hex_metric:
        return lib().or_ring_score_hex_m(n, rp, xp, C.c_float(sigma), qp, int(hex_metric))
    return getattr(lib(), {"tet": "or_ring_score_tet", "hex": "or_ring_score_hex", "tri": "or_ring_score_tri2d"}[kind])(
        n, rp, xp, C.c_float(sigma), qp)


def boundary_faces(kind: str, cells) -> np.ndarray:
    c, cp = _i(cells)
    k = 3 if kind == "tet" else 4
    per = 4 if kind == "tet" else 6
    out = np.empty(len(c) * per * k, dtype=np.int32)
    n = lib().or_boundary_faces(OR_TET if kind == "tet" else OR_HEX, len(c), cp, out.ctypes.data_as(_ip))
    return out[: n * k].reshape(n, k)


def color(mode: int, row_ptr, col_ind) -> np.ndarray:
    r, rp = _i(row_ptr)
    c, cp = _i(col_ind)
    n = len(r) - 1
    out = np.empty(max(n, 1), dtype=np.int32)
    lib().or_color(mode, n, rp, cp, out.ctypes.data_as(_ip))
    return out[:n]


_ELEM = {"tri": OR_TRI, "quad": OR_QUAD, "tet": OR_TET, "hex": OR_HEX}


def _arr(ptr, n, dtype=np.int32):
    if n <= 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class OracleMesh:
    """Owns an or_mesh built from host arrays; exposes numpy copies."""

    def __init__(self, kind: str, xyz, cells, coloring: int = OR_COLOR_GREEDY, feature_angle: float = 20.0):
        self.kind = kind
        x, xp = _f(xyz)
        c, cp = _i(cells)
        prm = BuildParams(coloring, feature_angle)
        out = C.POINTER(OrMesh)()
        self.L = lib()
        rc = self.L.or_mesh_build(_ELEM[kind], xp, len(x), cp, len(c), C.byref(prm), C.byref(out))
        if not out:
            raise ValueError(f"oracle mesh build failed: {rc}")
        self.ptr = out
        self.rc = rc
        # the hex metric of every search and quality call (OR_HEX_SHAPE9 or OR_HEX_CORNER8)
        self.hex_metric = OR_HEX_SHAPE9

    def __del__(self):
        if getattr(self, "surf", None):
            self.L.or_surface_free(self.surf)
            self.surf = None
        if getattr(self, "ptr", None):
            self.L.or_mesh_free(self.ptr)
            self.ptr = None

    @property
    def m(self) -> OrMesh:
        return self.ptr.contents

    def positions(self) -> np.ndarray:
        m = self.m
        return _arr(m.xyz, 3 * m.nV, np.float32).reshape(-1, 3)

    def set_positions(self, xyz):
        m = self.m
        a = np.ascontiguousarray(xyz, dtype=np.float32).ravel()
        C.memmove(m.xyz, a.ctypes.data, a.nbytes)

    def cells(self):
        m = self.m
        return _arr(m.cells, m.npe * m.nE).reshape(-1, m.npe)

    def new_of_old(self):
        return _arr(self.m.new_of_old, self.m.nV)

    def offsets(self, which: str):
        m = self.m
        if which == "inner":
            return _arr(m.colOffInner, m.nColorsInner + 1)
        if which == "free":
            return _arr(m.colOffFree, m.nColorsFree + 1)
        return _arr(m.colOffFeature, m.nColorsFeature + 1)

    def faces(self):
        m = self.m
        return _arr(m.faces, m.fpe * m.nFaces).reshape(-1, m.fpe)

    def ring(self):
        m = self.m
        row = _arr(m.ringRow, m.nV + 1)
        return row, _arr(m.ringVal, m.ringStride * int(row[-1])).reshape(-1, m.ringStride)

    def halfedges(self):
        m = self.m
        row = _arr(m.heRow, m.nSurf + 1)
        n = m.nHE
        return row, np.stack([_arr(m.heTarget, n), _arr(m.heOpp, n), _arr(m.heFace, n), _arr(m.heNext, n)], axis=1)

    def flags(self):
        m = self.m
        return _arr(m.isBoundary1D, m.nSurf, np.uint8), _arr(m.isFeature, m.nSurf, np.uint8)

    def iterate(self, nq=8, depth=3, grid_scale=0.5, shrink=0.0, extent=None, fixed_boundary=1, nthreads=0,
                max_per_class=0) -> int:
        """one iteration; max_per_class > 0 searches only the first vertices
        of each colour class (a timing sample).  Returns the searches run."""
        if extent is None:
            extent = OR_EXTENT_MIN if self.kind == "hex" else OR_EXTENT_MAX
        p = DmoParams(nq, depth, grid_scale, shrink, extent, fixed_boundary, nthreads, max_per_class, self.hex_metric)
        return self.L.or_dmo_iteration(self.ptr, C.byref(p))

    def search_vertex(self, vid, xyz, nq=8, depth=3, grid_scale=0.5, shrink=0.0, extent=None):
        if extent is None:
            extent = OR_EXTENT_MIN if self.kind == "hex" else OR_EXTENT_MAX
        p = DmoParams(nq, depth, grid_scale, shrink, extent, 1, 1, 0, self.hex_metric)
        a = np.ascontiguousarray(xyz, dtype=np.float32).copy()
        moved = self.L.or_dmo_vertex(self.ptr, C.byref(p), vid, a.ctypes.data_as(_fp))
        return moved, a

    def learned_window(self, vid, xyz, c, depth_l, frac, nq=8, depth=3, grid_scale=0.5):
        """the learned proposal's window search of inner vertex vid around c
        on the positions xyz: (status, position); status 1 moved, 0 stays,
        -1 falls back to the full search"""
        extent = OR_EXTENT_MIN if self.kind == "hex" else OR_EXTENT_MAX
        p = DmoParams(nq, depth, grid_scale, 0.0, extent, 1, 1, 0, self.hex_metric)
        a = np.ascontiguousarray(xyz, dtype=np.float32)
        cc = np.ascontiguousarray(c, dtype=np.float32).reshape(3)
        out = np.zeros(3, np.float32)
        st = self.L.or_learned_window(self.ptr, C.byref(p), vid, a.ctypes.data_as(_fp), cc.ctypes.data_as(_fp),
                                      depth_l, C.c_float(frac), out.ctypes.data_as(_fp))
        return st, out

    def surface_init(self, ring: int = 1):
        """Local surfaces etc. from the current positions (optimiser init);
        ring 2: quadric fits over the 2-ring"""
        if getattr(self, "surf", None):
            self.L.or_surface_free(self.surf)
        self.surf = self.L.or_surface_init_ring(self.ptr, int(ring))
        self.last_min = self.min_quality()

    def iterate_surface(self, nq=8, depth=3, grid_scale=0.5, shrink=0.0, extent=None, nthreads=0):
        """Inner pass, surface 2D / feature 1D passes gated on the last minimum
        quality, NN update (reference doIteration); returns (2D, 1D) counts."""
        if extent is None:
            extent = OR_EXTENT_MIN if self.kind == "hex" else OR_EXTENT_MAX
        p = DmoParams(nq, depth, grid_scale, shrink, extent, 0, nthreads, 0, self.hex_metric)
        a, b = C.c_longlong(), C.c_longlong()
        self.L.or_dmo_iteration_surface(self.ptr, self.surf, C.byref(p), C.c_float(self.last_min), C.byref(a), C.byref(b))
        self.last_min = self.min_quality()
        return a.value, b.value

    def iterate_surface_mesh(self, nq=8, depth=3, grid_scale=0.5, shrink=0.0, fixed_boundary=0, nthreads=0):
        """Surface mesh (tri / quad) on the surface path: 2D passes, 1D passes,
        NN and normal updates (reference DMOTriClass / DMOQuadClass
        doIteration); returns (2D, 1D) counts."""
        p = DmoParams(nq, depth, grid_scale, shrink, OR_EXTENT_MAX, fixed_boundary, nthreads)
        a, b = C.c_longlong(), C.c_longlong()
        self.L.or_dmo_iteration_surface_mesh(self.ptr, self.surf, C.byref(p), C.byref(a), C.byref(b))
        return a.value, b.value

    def surface_mesh_qualities(self) -> np.ndarray:
        out = np.empty(self.m.nE, dtype=np.float32)
        self.L.or_surface_mesh_qualities(self.ptr, out.ctypes.data_as(_fp))
        return out

    def surface_mesh_min_quality(self) -> float:
        return self.L.or_surface_mesh_min_quality(self.ptr)

    def laplace(self, n_iter: int):
        self.L.or_laplace(self.ptr, int(n_iter))

    def surface_points(self, vid: int, nu: int, nv: int, mode: int, sid: int = 0, grid_scale: float = 0.5):
        out = np.zeros((nu * nv, 3), dtype=np.float32)
        self.L.or_surface_points(self.ptr, self.surf, vid, nu, nv, mode, sid, C.c_float(grid_scale),
                                out.ctypes.data_as(_fp))
        return out

    def surface_get(self, what: int):
        n = self.m.nSurf
        if what == 0:
            out = np.empty(n, np.int32)
        elif what == 1:
            out = np.empty((n, 16), np.float32)
        elif what == 2:
            out = np.empty(n, np.uint8)
        elif what == 4:
            out = np.empty((n, 3), np.float32)
        else:
            out = np.empty((n, 10), np.float32)
        self.L.or_surface_get(self.surf, what, out.ctypes.data_as(C.c_void_p))
        return out

    def sectors(self):
        """(table int[nFeature, 20], row offsets int[nFeature + 1], fits
        float[nSectors, 16])"""
        nF = self.m.nFeature
        table = np.empty((max(nF, 1), 20), np.int32)
        frow = np.empty(nF + 1, np.int32)
        self.L.or_surface_get_sectors(self.surf, table.ctypes.data_as(C.c_void_p), frow.ctypes.data_as(C.c_void_p), None)
        fits = np.empty((max(int(frow[-1]), 1), 16), np.float32)
        self.L.or_surface_get_sectors(self.surf, None, None, fits.ctypes.data_as(C.c_void_p))
        return table[:nF], frow, fits[: int(frow[-1])]

    def surface_search_vertex(self, vid: int, dim: int, nq=8, depth=3, grid_scale=0.5, shrink=0.0):
        """One surface vertex's 2D (dim 2) or 1D (dim 1) search on the current
        positions; returns (written, final position, final local (u, v))."""
        extent = OR_EXTENT_MIN if self.kind == "hex" else OR_EXTENT_MAX
        p = DmoParams(nq, depth, grid_scale, shrink, extent, 0, 1, 0, self.hex_metric)
        a = self.positions().copy()
        uv = np.zeros(2, np.float32)
        w = self.L.or_surface_search_vertex(self.ptr, self.surf, C.byref(p), vid, dim, a.ctypes.data_as(_fp),
                                            uv.ctypes.data_as(_fp))
        return w, a[vid], uv

    def surface_error(self) -> int:
        """bit 0: a projection found no feature sector; bit 1: a feature vertex
        has neighbours that belong to no sector; bit 2: a ring-2 fit would
        hold more than 64 points"""
        return self.L.or_surface_error(self.surf)

    def surface_project(self, vid: int, u: float, v: float) -> np.ndarray:
        out = np.zeros(3, np.float32)
        self.L.or_surface_project(self.ptr, self.surf, vid, C.c_float(u), C.c_float(v), out.ctypes.data_as(_fp))
        return out

    def surface_update_nn(self):
        self.L.or_surface_update_nn(self.ptr, self.surf)

    def min_quality(self) -> float:
        return self.L.or_min_quality_m(self.ptr, self.hex_metric)

    def quality64(self):
        a, b = C.c_double(), C.c_double()
        self.L.or_quality64_m(self.ptr, self.hex_metric, C.byref(a), C.byref(b))
        return a.value, b.value

    def element_qualities(self) -> np.ndarray:
        out = np.empty(self.m.nE, dtype=np.float32)
        self.L.or_element_qualities_m(self.ptr, self.hex_metric, out.ctypes.data_as(_fp))
        return out


def search_ring(kind: str, ring, x0, nq=8, depth=3, grid_scale=0.5, shrink=0.0, extent=None, hex_metric=0):
    """The oracle's vertex search on an explicit ring (tet: n x 3 points,
    hex: n x 7 points, tri: a closed fan of n points); returns (written,
    final position)."""
    if extent is None:
        extent = OR_EXTENT_MIN if kind == "hex" else OR_EXTENT_MAX
    r = np.ascontiguousarray(ring, dtype=np.float32).reshape(-1, 3)
    per = {"tet": 3, "hex": 7, "tri": 1}[kind]
    x, xp = _f(np.asarray(x0, dtype=np.float32).reshape(3))
    out = np.zeros(3, np.float32)
    p = DmoParams(nq, depth, grid_scale, shrink, extent, 1, 1, 0, hex_metric)
    w = lib().or_search_ring(_ELEM[kind], r.shape[0] // per, r.ctypes.data_as(_fp), xp, C.byref(p),
                             out.ctypes.data_as(_fp))
    return w, out


def fit_local_surface(pts, normal, refp) -> np.ndarray:
    """Quadric fit of pts (pts[0] = the centre) in the frame of (normal,
    refp - centre): 16 floats p0, e0, e1, e2, a, b, c, numPoints."""
    p = np.ascontiguousarray(pts, dtype=np.float32)
    n = np.ascontiguousarray(normal, dtype=np.float32)
    r = np.ascontiguousarray(refp, dtype=np.float32)
    out = np.zeros(16, dtype=np.float32)
    lib().or_fit_local_surface(p.ctypes.data_as(_fp), p.shape[0], n.ctypes.data_as(_fp), r.ctypes.data_as(_fp),
                               out.ctypes.data_as(_fp))
    return out


def lu_solve3(A, b) -> np.ndarray:
    Acm = np.ascontiguousarray(np.asarray(A, dtype=np.float32).T)   # column-major
    bb = np.ascontiguousarray(b, dtype=np.float32)
    x = np.zeros(3, dtype=np.float32)
    lib().or_lu_solve3(Acm.ctypes.data_as(_fp), bb.ctypes.data_as(_fp), x.ctypes.data_as(_fp))
    return x

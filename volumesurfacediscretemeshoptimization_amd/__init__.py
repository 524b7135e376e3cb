"""MI355X-native discrete mesh optimisation (DMO).

Python mirror of the reference's mesh/optimiser interface over the C ABI of
libvsdmo.so (include/vsdmo.h):

* ``DMOMeshTriFactory.create`` / ``DMOMeshQuadFactory.create`` /
  ``DMOMeshTetFactory.create`` / ``DMOMeshHexFactory.create`` (reference
  VSDMO/src/mesh/MeshFactory.h:17-41) build the device mesh: boundary faces,
  renumbering, colouring, one-rings, surface halfedges.
* ``DMOTetClass`` / ``DMOHexClass`` / ``DMOTriFlatClass`` / ``DMOTriClass`` /
  ``DMOQuadClass`` (VSDMO/src/dmo/DMOTet.h:24, DMOHex.h:18, DMOTriFlat.h:25,
  DMOTri.h:24, DMOQuad.h:18) with the virtual interface of ``DMOBaseClass``
  (VSDMO/src/dmo/DMOBase.h:23-88): ``optimize``, ``doIteration``,
  ``isDone``, ``getQuality``, ``getQualityHistogram``,
  ``getElementQualities``, ``displayQualityGPU``.
* ``laplaceSmoothing(mesh, n_iter)`` (VSDMO/src/dmo/Laplace.h:10).

All arithmetic runs in the HIP kernels of libvsdmo.so; this module only moves
host arrays across the ABI.  A missing library raises; there is no fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import VsdmoError, check, load  # noqa: F401

MEAN_RATIO = 0
# hex metrics: the reference's 9 sub-tets, or the 8 corner tets alone
HEX_SHAPE9, HEX_CORNER8 = _lib.VSDMO_HEX_SHAPE9, _lib.VSDMO_HEX_CORNER8

__all__ = ["read_mesh_file", "write_mesh_file", "DMOMeshTriFactory", "DMOMeshQuadFactory", "DMOMeshTetFactory",
           "DMOMeshHexFactory", "DMOMesh", "DMOTetClass", "DMOHexClass", "DMOTriFlatClass", "DMOTriClass", "DMOQuadClass",
           "laplaceSmoothing", "DMOBatch", "DMOMLP", "DMODataset", "VsdmoError", "MEAN_RATIO", "HEX_SHAPE9",
           "HEX_CORNER8", "load"]


def _fptr(a):
    return a.ctypes.data_as(_lib._fp)


def _iptr(a):
    return a.ctypes.data_as(_lib._ip)


class DMOMesh:
    """Device mesh (owns the vsdmo_mesh handle)."""

    NPE = {_lib.VSDMO_TRI: 3, _lib.VSDMO_QUAD: 4, _lib.VSDMO_TET: 4, _lib.VSDMO_HEX: 8}

    def __init__(self, elem: int, xyz, cells, coloring: int = _lib.VSDMO_COLOR_GREEDY, feature_angle: float = 20.0,
                 coloring_polls: int = 0):
        lib = load()
        self._lib = lib
        self.elem = elem
        xyz = np.ascontiguousarray(xyz, dtype=np.float32)
        cells = np.ascontiguousarray(cells, dtype=np.int32)
        if xyz.ndim != 2 or xyz.shape[1] != 3:
            raise ValueError("xyz must have shape (nV, 3)")
        npe = self.NPE[elem]
        if cells.ndim != 2 or cells.shape[1] != npe:
            raise ValueError(f"cells must have shape (nE, {npe})")
        prm = _lib.MeshParams(coloring, feature_angle, coloring_polls)
        h = C.c_void_p()
        check(lib.vsdmo_mesh_create(elem, _fptr(xyz), xyz.shape[0], _iptr(cells), cells.shape[0], None,
                                    C.byref(prm), C.byref(h)), "vsdmo_mesh_create")
        self._h = h
        self.npe = npe

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            check(self._lib.vsdmo_mesh_destroy(self._h), "vsdmo_mesh_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def write(self, path: str, original_order: bool = True):
        """Writes the current mesh (OFF / TetGen / MEDIT by element type)."""
        check(self._lib.vsdmo_mesh_write(self._h, path.encode(), int(original_order)), "vsdmo_mesh_write")

    def counts(self) -> dict:
        c = _lib.Counts()
        check(self._lib.vsdmo_mesh_counts(self._h, C.byref(c)), "vsdmo_mesh_counts")
        return {name: getattr(c, name) for name, _ in _lib.Counts._fields_}

    @property
    def nVertices(self) -> int:
        return self.counts()["n_vertices"]

    def positions(self, original_order: bool = True) -> np.ndarray:
        n = self.nVertices
        out = np.empty((n, 3), dtype=np.float32)
        check(self._lib.vsdmo_mesh_get_positions(self._h, _fptr(out), int(original_order)), "get_positions")
        return out

    def set_positions(self, xyz, original_order: bool = True):
        a = np.ascontiguousarray(xyz, dtype=np.float32)
        check(self._lib.vsdmo_mesh_set_positions(self._h, _fptr(a), int(original_order)), "set_positions")

    def permutation(self) -> np.ndarray:
        out = np.empty(self.nVertices, dtype=np.int32)
        check(self._lib.vsdmo_mesh_get_permutation(self._h, _iptr(out)), "get_permutation")
        return out

    def cells(self) -> np.ndarray:
        c = self.counts()
        out = np.empty((c["n_elements"], self.npe), dtype=np.int32)
        check(self._lib.vsdmo_mesh_get_cells(self._h, _iptr(out)), "get_cells")
        return out

    def color_offsets(self, which: str) -> np.ndarray:
        c = self.counts()
        idx = {"inner": 0, "free": 1, "feature": 2}[which]
        n = c[{"inner": "n_colors_inner", "free": "n_colors_free", "feature": "n_colors_feature"}[which]]
        out = np.empty(n + 1, dtype=np.int32)
        check(self._lib.vsdmo_mesh_get_coloring(self._h, idx, _iptr(out)), "get_coloring")
        return out

    def check_coloring(self, offsets):
        """Checks a class partition [offsets[c], offsets[c+1]) of the current
        numbering: VsdmoError (VSDMO_TOPOLOGY) when an element holds two
        vertices of one class."""
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        check(self._lib.vsdmo_mesh_check_coloring(self._h, _iptr(off), len(off) - 1), "check_coloring")

    def colors(self) -> np.ndarray:
        out = np.empty(self.nVertices, dtype=np.int32)
        check(self._lib.vsdmo_mesh_get_colors(self._h, _iptr(out)), "get_colors")
        return out

    def faces(self) -> np.ndarray:
        c = self.counts()
        k = 3 if self.elem in (_lib.VSDMO_TET, _lib.VSDMO_TRI) else 4
        out = np.empty((max(c["n_faces"], 0), k), dtype=np.int32)
        if c["n_faces"]:
            check(self._lib.vsdmo_mesh_get_faces(self._h, _iptr(out)), "get_faces")
        return out

    def onering(self):
        n = self.nVertices
        row = np.empty(n + 1, dtype=np.int32)
        check(self._lib.vsdmo_mesh_get_onering(self._h, _iptr(row), None), "get_onering")
        st = 7 if self.elem == _lib.VSDMO_HEX else 3
        vals = np.empty((max(int(row[-1]), 1), st), dtype=np.int32)
        check(self._lib.vsdmo_mesh_get_onering(self._h, _iptr(row), _iptr(vals)), "get_onering")
        return row, vals[: int(row[-1])]

    def halfedges(self):
        c = self.counts()
        row = np.empty(c["n_surface"] + 1, dtype=np.int32)
        n = max(c["n_halfedges"], 1)
        t, o, f, nx = (np.empty(n, dtype=np.int32) for _ in range(4))
        check(self._lib.vsdmo_mesh_get_halfedges(self._h, _iptr(row), _iptr(t), _iptr(o), _iptr(f), _iptr(nx)),
              "get_halfedges")
        k = c["n_halfedges"]
        return row, np.stack([t[:k], o[:k], f[:k], nx[:k]], axis=1)

    def flags(self):
        c = self.counts()
        n = max(c["n_surface"], 1)
        b = np.zeros(n, dtype=np.uint8)
        f = np.zeros(n, dtype=np.uint8)
        check(self._lib.vsdmo_mesh_get_flags(self._h, b.ctypes.data_as(_lib._u8p), f.ctypes.data_as(_lib._u8p)),
              "get_flags")
        return b[: c["n_surface"]], f[: c["n_surface"]]


def read_mesh_file(elem: int, path: str):
    """(xyz, cells) of a mesh file: OFF (flat triangles), TetGen base path
    (tets, <path>.node / <path>.ele) or MEDIT .mesh (hexes, internal order)."""
    lib = load()
    nv, nc = C.c_int32(), C.c_int32()
    check(lib.vsdmo_mesh_file_read(elem, path.encode(), C.byref(nv), C.byref(nc), None, None), "vsdmo_mesh_file_read")
    xyz = np.empty((nv.value, 3), dtype=np.float32)
    cells = np.empty((nc.value, DMOMesh.NPE[elem]), dtype=np.int32)
    check(lib.vsdmo_mesh_file_read(elem, path.encode(), C.byref(nv), C.byref(nc), _fptr(xyz), _iptr(cells)),
          "vsdmo_mesh_file_read")
    return xyz, cells


def write_mesh_file(elem: int, path: str, xyz, cells):
    lib = load()
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    check(lib.vsdmo_mesh_file_write(elem, path.encode(), _fptr(xyz), xyz.shape[0], _iptr(cells), cells.shape[0]),
          "vsdmo_mesh_file_write")


class _Factory:
    ELEM = None

    @classmethod
    def create(cls, xyz_or_path, cells=None, **kw):
        """create(xyz, cells) from host arrays, or create(path) from a mesh
        file (reference DMOMesh*Factory::create(std::string); None when the
        file cannot be read, like the reference's nullptr)."""
        if isinstance(xyz_or_path, (str, bytes)):
            path = xyz_or_path.decode() if isinstance(xyz_or_path, bytes) else xyz_or_path
            try:
                xyz, cells = read_mesh_file(cls.ELEM, path)
            except VsdmoError:
                return None
            return DMOMesh(cls.ELEM, xyz, cells, **kw)
        return DMOMesh(cls.ELEM, xyz_or_path, cells, **kw)


class DMOMeshTriFactory(_Factory):
    ELEM = _lib.VSDMO_TRI


class DMOMeshQuadFactory(_Factory):
    ELEM = _lib.VSDMO_QUAD


class DMOMeshTetFactory(_Factory):
    ELEM = _lib.VSDMO_TET


class DMOMeshHexFactory(_Factory):
    ELEM = _lib.VSDMO_HEX


class _DMOBase:
    """Reference DMOBaseClass (VSDMO/src/dmo/DMOBase.h:23-88)."""

    SURFACE_PATH = _lib.VSDMO_SURFACE_PATH_AUTO

    def __init__(self, mesh: DMOMesh, qualityCriterium: int = MEAN_RATIO, gridScale: float = 0.5,
                 n_iter: int = 30, nq: int = 8, depth: int = 3, shrink: float = 0.0,
                 extent_mode: int = _lib.VSDMO_EXTENT_AUTO, fixed_boundary: bool = False,
                 quality_guard: bool = True, guard_tolerance: float = 1e-5, waves_per_vertex: int = 1,
                 surface_path: int = None, hex_metric: int = _lib.VSDMO_HEX_SHAPE9, surface_ring: int = 1):
        lib = load()
        self._lib = lib
        self.mesh = mesh
        p = _lib.Params()
        lib.vsdmo_params_default(C.byref(p))
        p.nq, p.depth, p.shrink, p.extent_mode = nq, depth, shrink, extent_mode
        p.fixed_boundary, p.quality_guard, p.guard_tolerance = int(fixed_boundary), int(quality_guard), guard_tolerance
        p.waves_per_vertex = int(waves_per_vertex)
        p.surface_path = self.SURFACE_PATH if surface_path is None else int(surface_path)
        p.hex_metric, p.surface_ring = int(hex_metric), int(surface_ring)
        h = C.c_void_p()
        check(lib.vsdmo_opt_create(mesh.handle, qualityCriterium, gridScale, n_iter, C.byref(p), C.byref(h)),
              "vsdmo_opt_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            check(self._lib.vsdmo_opt_destroy(self._h), "vsdmo_opt_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def doIteration(self):
        check(self._lib.vsdmo_opt_iterate(self._h, 1), "vsdmo_opt_iterate")

    def iterate(self, n: int):
        check(self._lib.vsdmo_opt_iterate(self._h, int(n)), "vsdmo_opt_iterate")

    def optimize(self):
        check(self._lib.vsdmo_opt_optimize(self._h), "vsdmo_opt_optimize")

    def isDone(self) -> bool:
        return bool(self._lib.vsdmo_opt_is_done(self._h))

    def getQuality(self) -> float:
        q = C.c_float()
        check(self._lib.vsdmo_opt_min_quality(self._h, C.byref(q)), "vsdmo_opt_min_quality")
        return q.value

    def getMeanQuality(self) -> float:
        q = C.c_double()
        check(self._lib.vsdmo_opt_mean_quality(self._h, C.byref(q)), "vsdmo_opt_mean_quality")
        return q.value

    def getQualityHistogram(self, n_cols: int = 10) -> np.ndarray:
        out = np.zeros(n_cols, dtype=np.int32)
        check(self._lib.vsdmo_opt_histogram(self._h, n_cols, _iptr(out)), "vsdmo_opt_histogram")
        return out

    def getElementQualities(self) -> np.ndarray:
        n = self.mesh.counts()["n_elements"]
        out = np.empty(n, dtype=np.float32)
        check(self._lib.vsdmo_opt_element_qualities(self._h, _fptr(out)), "vsdmo_opt_element_qualities")
        return out

    def displayQualityGPU(self, n_cols: int = 10):
        h = self.getQualityHistogram(n_cols)
        print("interval " + " ".join(f"{i / n_cols:.2f}" for i in range(n_cols)))
        print("amount   " + " ".join(str(int(v)) for v in h))
        print(f"q_min = {self.getQuality()}")

    def set_timing(self, on: bool):
        check(self._lib.vsdmo_opt_set_timing(self._h, int(on)), "vsdmo_opt_set_timing")

    def set_failure_dump(self, path):
        """Write the mesh to path (its own format, renumbered order) when the
        quality guard fails (reference updateCurrentQualities's
        res/fail_debug dump); None turns it off."""
        check(self._lib.vsdmo_opt_set_failure_dump(self._h, path.encode() if path else None),
              "vsdmo_opt_set_failure_dump")

    def stats(self) -> dict:
        s = _lib.Stats()
        check(self._lib.vsdmo_opt_stats(self._h, C.byref(s)), "vsdmo_opt_stats")
        return {name: getattr(s, name) for name, _ in _lib.Stats._fields_}


    def getEstimateLocalSurfacePoints(self, vid: int, nu: int, nv: int) -> np.ndarray:
        """nu*nv points of the blended surface projection around surface
        vertex vid (optimiser numbering)."""
        out = np.empty((nu * nv, 3), dtype=np.float32)
        check(self._lib.vsdmo_opt_surface_points(self._h, int(vid), int(nu), int(nv), 0, 0, _fptr(out)),
              "vsdmo_opt_surface_points")
        return out

    def getLocalSurfacePoints(self, vid: int, nu: int, nv: int, featureSid: int = 0) -> np.ndarray:
        """nu*nv points of the local fit of vid's nearest neighbour (its
        sector fit featureSid for a feature vertex)."""
        out = np.empty((nu * nv, 3), dtype=np.float32)
        check(self._lib.vsdmo_opt_surface_points(self._h, int(vid), int(nu), int(nv), 1, int(featureSid), _fptr(out)),
              "vsdmo_opt_surface_points")
        return out

    # default proposal windows (levels, fraction of the plain first-level
    # window), chosen with the deferred fallback on 40^3 tet and 48^3 hex cubes
    LEARN_WINDOW = {_lib.VSDMO_TET: (1, 0.025), _lib.VSDMO_HEX: (1, 2.0)}

    def set_proposal(self, mlp=None, mode: int = 1, depth: int = None, window_frac: float = None,
                     fallback: str = None):
        """Learned proposal for the inner searches: mode 1 = network, 2 = the
        vertex itself (same windowed search and fallback, no network), 3 = the
        network, then (where its window finds nothing better) a window a
        quarter of its size around the vertex itself before the full search,
        0 = off.
        The window is `depth` levels of 4^3 candidates on `window_frac` times
        the plain first-level window (defaults per element type above).
        fallback: "immediate" (the full search of a vertex whose window finds
        nothing better runs in the same class pass) or "deferred" (it runs in
        the class's next pass, beside that pass's windows); None keeps the
        current policy (initially immediate)."""
        d0, f0 = self.LEARN_WINDOW.get(self.mesh.elem, (1, 0.025))
        depth = d0 if depth is None else depth
        window_frac = f0 if window_frac is None else window_frac
        if fallback is not None:
            self.set_fallback(fallback)
        h = mlp.handle if mlp is not None else None
        check(self._lib.vsdmo_opt_set_proposal(self._h, h, int(mode), int(depth), float(window_frac)),
              "vsdmo_opt_set_proposal")
        self._mlp = mlp if mode in (1, 3) else None   # the network is destroyed after the optimiser

    FALLBACK = {"immediate": 0, "deferred": 1}

    def set_fallback(self, policy: str):
        """Fallback policy of the learned proposal ("immediate" / "deferred");
        clears pending deferred searches."""
        if policy not in self.FALLBACK:
            raise ValueError(f"fallback must be one of {sorted(self.FALLBACK)}")
        check(self._lib.vsdmo_opt_set_fallback(self._h, self.FALLBACK[policy]), "vsdmo_opt_set_fallback")

    def record(self, dataset, n: int):
        """n plain iterations, recording one training pair per inner search."""
        check(self._lib.vsdmo_opt_record(self._h, dataset.handle, int(n)), "vsdmo_opt_record")


class DMOMLP:
    """96-128-128-3 GELU MLP with BF16 MFMA kernels (the learned proposal)."""

    def __init__(self, seed: int = 0):
        lib = load()
        self._lib = lib
        h = C.c_void_p()
        check(lib.vsdmo_mlp_create(C.c_uint64(seed), C.byref(h)), "vsdmo_mlp_create")
        self._h = h
        self.n_params = lib.vsdmo_mlp_num_params()

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            check(self._lib.vsdmo_mlp_destroy(self._h), "vsdmo_mlp_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.empty(self.n_params, dtype=np.float32)
        check(self._lib.vsdmo_mlp_get_params(self._h, _fptr(out)), "vsdmo_mlp_get_params")
        return out

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        assert a.size == self.n_params
        check(self._lib.vsdmo_mlp_set_params(self._h, _fptr(a)), "vsdmo_mlp_set_params")

    def forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        n = X.shape[0]
        Y = np.empty((n, 3), dtype=np.float32)
        check(self._lib.vsdmo_mlp_forward(self._h, _fptr(X), n, _fptr(Y)), "vsdmo_mlp_forward")
        return Y

    def gradients(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.empty(self.n_params, dtype=np.float32)
        loss = C.c_double()
        check(self._lib.vsdmo_mlp_gradients(self._h, _fptr(X), _fptr(T), X.shape[0], C.byref(loss), _fptr(g)),
              "vsdmo_mlp_gradients")
        return loss.value, g

    def train(self, dataset, epochs: int = 20, batch: int = 4096, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        out = np.zeros(epochs, dtype=np.float64)
        check(self._lib.vsdmo_mlp_train(self._h, dataset.handle, int(epochs), int(batch), float(lr), C.c_uint64(seed),
                                        out.ctypes.data_as(C.POINTER(C.c_double))), "vsdmo_mlp_train")
        return out


class DMODataset:
    """Device-resident training pairs (one-ring features, displacement / L)."""

    def __init__(self, capacity: int):
        lib = load()
        self._lib = lib
        h = C.c_void_p()
        check(lib.vsdmo_dataset_create(C.c_int64(capacity), C.byref(h)), "vsdmo_dataset_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            check(self._lib.vsdmo_dataset_destroy(self._h), "vsdmo_dataset_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        n = C.c_int64()
        check(self._lib.vsdmo_dataset_size(self._h, C.byref(n)), "vsdmo_dataset_size")
        return n.value

    def get(self, start: int, n: int):
        X = np.empty((n, 96), dtype=np.float32)
        T = np.empty((n, 3), dtype=np.float32)
        check(self._lib.vsdmo_dataset_get(self._h, C.c_int64(start), int(n), _fptr(X), _fptr(T)), "vsdmo_dataset_get")
        return X, T


class DMOTetClass(_DMOBase):
    pass


class DMOHexClass(_DMOBase):
    pass


class DMOTriFlatClass(_DMOBase):
    """Flat triangle meshes: 2D search in the plane (DMOTriFlat.h:25)."""
    SURFACE_PATH = _lib.VSDMO_SURFACE_PATH_FLAT


class DMOTriClass(_DMOBase):
    """Triangle surface meshes: free vertices move on their local surfaces,
    feature vertices along their feature curves (DMOTri.h:24)."""
    SURFACE_PATH = _lib.VSDMO_SURFACE_PATH_SURFACE


class DMOQuadClass(_DMOBase):
    """Quad surface meshes (DMOQuad.h:18)."""
    SURFACE_PATH = _lib.VSDMO_SURFACE_PATH_SURFACE


def laplaceSmoothing(mesh: DMOMesh, n_iter: int):
    """Laplace smoothing of the free vertices of a triangle / quad mesh,
    colour class by colour class (reference dmo/Laplace.cu:30-41)."""
    check(mesh._lib.vsdmo_mesh_laplace(mesh.handle, int(n_iter)), "vsdmo_mesh_laplace")


class DMOBatch:
    """Batched DMO over many meshes (no counterpart in the reference): colour
    round r sweeps colour r of every mesh in one persistent launch, one
    iteration is replayed as a hipGraph.  Each mesh gets the positions one
    DMO*Class per mesh (boundary fixed) would give."""

    def __init__(self, meshes, nq: int = 8, depth: int = 3, waves_per_vertex: int = 1, gridScale: float = 0.5,
                 use_graph: bool = True, shuffle_seed: int = 0, quality_guard: bool = True,
                 guard_tolerance: float = 1e-5, persistent: bool = True, hex_metric: int = _lib.VSDMO_HEX_SHAPE9):
        lib = load()
        self._lib = lib
        self.meshes = list(meshes)
        p = _lib.BatchParams()
        lib.vsdmo_batch_params_default(C.byref(p))
        p.nq, p.depth, p.waves_per_vertex, p.grid_scale = nq, depth, waves_per_vertex, gridScale
        p.use_graph, p.shuffle_seed = int(use_graph), int(shuffle_seed)
        p.quality_guard, p.guard_tolerance = int(quality_guard), guard_tolerance
        p.persistent = int(persistent)
        p.hex_metric = int(hex_metric)
        arr = (C.c_void_p * len(self.meshes))(*[m.handle.value for m in self.meshes])
        h = C.c_void_p()
        check(lib.vsdmo_batch_create(arr, len(self.meshes), C.byref(p), C.byref(h)), "vsdmo_batch_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            check(self._lib.vsdmo_batch_destroy(self._h), "vsdmo_batch_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def iterate(self, n: int = 1):
        check(self._lib.vsdmo_batch_iterate(self._h, int(n)), "vsdmo_batch_iterate")

    def min_quality(self) -> np.ndarray:
        out = np.empty(len(self.meshes), dtype=np.float32)
        check(self._lib.vsdmo_batch_min_quality(self._h, _fptr(out)), "vsdmo_batch_min_quality")
        return out

    def mean_quality(self) -> np.ndarray:
        out = np.empty(len(self.meshes), dtype=np.float64)
        check(self._lib.vsdmo_batch_mean_quality(self._h, out.ctypes.data_as(C.POINTER(C.c_double))),
              "vsdmo_batch_mean_quality")
        return out

    def positions(self, i: int) -> np.ndarray:
        """Positions of mesh i in the optimiser numbering."""
        n = self.meshes[i].nVertices
        out = np.empty((n, 3), dtype=np.float32)
        check(self._lib.vsdmo_batch_get_positions(self._h, int(i), _fptr(out)), "vsdmo_batch_get_positions")
        return out

    def writeback(self):
        check(self._lib.vsdmo_batch_writeback(self._h), "vsdmo_batch_writeback")

    def restore(self):
        check(self._lib.vsdmo_batch_restore(self._h), "vsdmo_batch_restore")

    def params(self) -> dict:
        p = _lib.BatchParams()
        check(self._lib.vsdmo_batch_get_params(self._h, C.byref(p)), "vsdmo_batch_get_params")
        return {name: getattr(p, name) for name, _ in _lib.BatchParams._fields_}

    def set_params(self, **kw):
        p = _lib.BatchParams()
        check(self._lib.vsdmo_batch_get_params(self._h, C.byref(p)), "vsdmo_batch_get_params")
        for k, v in kw.items():
            setattr(p, k, v)
        check(self._lib.vsdmo_batch_set_params(self._h, C.byref(p)), "vsdmo_batch_set_params")

    def autotune(self, iters: int = 5, tol: float = 1e-3, log_path: str = "") -> dict:
        p = _lib.BatchParams()
        check(self._lib.vsdmo_batch_autotune(self._h, int(iters), float(tol), log_path.encode(), C.byref(p)),
              "vsdmo_batch_autotune")
        return {name: getattr(p, name) for name, _ in _lib.BatchParams._fields_}

    def stats(self) -> dict:
        s = _lib.BatchStats()
        check(self._lib.vsdmo_batch_get_stats(self._h, C.byref(s)), "vsdmo_batch_get_stats")
        return {name: getattr(s, name) for name, _ in _lib.BatchStats._fields_}

    def set_proposal(self, mlp=None, mode: int = 1, depth: int = None, window_frac: float = None,
                     elem: int = _lib.VSDMO_TET):
        """Learned proposal for the vertices of the batch's meshes of type
        elem (VSDMO_TET or VSDMO_HEX; each type has its own network and
        window).  A vertex whose window finds nothing better takes the full
        search in the same round, or in its round of the next iteration
        after set_fallback("deferred").  The other meshes keep the plain
        search.  Each such mesh gets the positions of a DMOTetClass /
        DMOHexClass with fixed_boundary=True and the same proposal and
        policy.  Defaults: the type's window in DMOTetClass /
        DMOHexClass."""
        d0, f0 = _DMOBase.LEARN_WINDOW.get(elem, (1, 0.025))
        depth = d0 if depth is None else depth
        window_frac = f0 if window_frac is None else window_frac
        h = mlp.handle if mlp is not None else None
        check(self._lib.vsdmo_batch_set_proposal(self._h, int(elem), h, int(mode), int(depth), float(window_frac)),
              "vsdmo_batch_set_proposal")
        if not hasattr(self, "_mlps"):
            self._mlps = {}
        self._mlps[int(elem)] = mlp if mode in (1, 3) else None   # networks are destroyed after the batch

    def set_fallback(self, policy: str):
        """Fallback policy of the learned proposals ("immediate", the
        default, or "deferred"); clears pending deferred searches."""
        if policy not in _DMOBase.FALLBACK:
            raise ValueError(f"fallback must be one of {sorted(_DMOBase.FALLBACK)}")
        check(self._lib.vsdmo_batch_set_fallback(self._h, _DMOBase.FALLBACK[policy]), "vsdmo_batch_set_fallback")

    def learned_stats(self) -> dict:
        v, f = C.c_int64(), C.c_int64()
        check(self._lib.vsdmo_batch_learned_stats(self._h, C.byref(v), C.byref(f)), "vsdmo_batch_learned_stats")
        return {"learned_visits": v.value, "fallbacks": f.value}

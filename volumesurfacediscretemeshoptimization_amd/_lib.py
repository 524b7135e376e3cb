"""ctypes loader for libvsdmo.so (the HIP library) and its C ABI signatures.

The library must exist; there is no fallback.  Build it with
``python -m volumesurfacediscretemeshoptimization_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvsdmo.so")

_ip = C.POINTER(C.c_int32)
_fp = C.POINTER(C.c_float)
_u8p = C.POINTER(C.c_uint8)

VSDMO_OK, VSDMO_INVALID_ARG, VSDMO_HIP_ERROR, VSDMO_TOPOLOGY = 0, 1, 2, 3
VSDMO_QUALITY_DECREASED, VSDMO_LDS_OVERFLOW, VSDMO_NOT_SUPPORTED = 4, 5, 6
STATUS_NAMES = {0: "OK", 1: "INVALID_ARG", 2: "HIP_ERROR", 3: "TOPOLOGY", 4: "QUALITY_DECREASED",
                5: "LDS_OVERFLOW", 6: "NOT_SUPPORTED"}
VSDMO_TRI, VSDMO_QUAD, VSDMO_TET, VSDMO_HEX = 0, 1, 2, 3
VSDMO_COLOR_GREEDY, VSDMO_COLOR_JPL = 0, 1
VSDMO_EXTENT_AUTO, VSDMO_EXTENT_MAX, VSDMO_EXTENT_MIN = -1, 0, 1
VSDMO_SURFACE_PATH_AUTO, VSDMO_SURFACE_PATH_SURFACE, VSDMO_SURFACE_PATH_FLAT = 0, 1, 2
VSDMO_HEX_SHAPE9, VSDMO_HEX_CORNER8 = 0, 1


class MeshParams(C.Structure):
    _fields_ = [("coloring", C.c_int32), ("feature_angle_deg", C.c_float), ("coloring_polls", C.c_int32)]


class Params(C.Structure):
    _fields_ = [("nq", C.c_int32), ("depth", C.c_int32), ("shrink", C.c_float), ("extent_mode", C.c_int32),
                ("fixed_boundary", C.c_int32), ("quality_guard", C.c_int32), ("guard_tolerance", C.c_float),
                ("waves_per_vertex", C.c_int32), ("surface_path", C.c_int32), ("hex_metric", C.c_int32),
                ("surface_ring", C.c_int32)]


class Counts(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "n_vertices", "n_elements", "n_surface", "n_free", "n_feature", "n_inner",
        "n_colors_inner", "n_colors_free", "n_colors_feature",
        "n_faces", "n_halfedges", "max_ring", "max_halfedges", "flat")]


class Stats(C.Structure):
    _fields_ = [("searches_inner", C.c_int64), ("searches_surface", C.c_int64), ("searches_feature", C.c_int64),
                ("iterations", C.c_int32), ("min_quality", C.c_float), ("kernel_ms", C.c_double),
                ("kernel_launches", C.c_int64), ("searches_learned", C.c_int64), ("fallbacks", C.c_int64)]


class BatchParams(C.Structure):
    _fields_ = [("nq", C.c_int32), ("depth", C.c_int32), ("waves_per_vertex", C.c_int32), ("grid_scale", C.c_float),
                ("use_graph", C.c_int32), ("shuffle_seed", C.c_uint32), ("quality_guard", C.c_int32),
                ("guard_tolerance", C.c_float), ("persistent", C.c_int32), ("hex_metric", C.c_int32)]


class BatchStats(C.Structure):
    _fields_ = [("searches", C.c_int64), ("iterations", C.c_int32), ("rounds", C.c_int32), ("meshes", C.c_int32),
                ("items_per_iteration", C.c_int64), ("gpu_ms", C.c_double), ("max_grid", C.c_int32),
                ("total_vertices", C.c_int32)]


EXPORTS = {
    "vsdmo_mesh_params_default": (None, [C.POINTER(MeshParams)]),
    "vsdmo_params_default": (None, [C.POINTER(Params)]),
    "vsdmo_mesh_create": (C.c_int, [C.c_int, _fp, C.c_int32, _ip, C.c_int32, _u8p, C.POINTER(MeshParams),
                                    C.POINTER(C.c_void_p)]),
    "vsdmo_mesh_destroy": (C.c_int, [C.c_void_p]),
    "vsdmo_mesh_counts": (C.c_int, [C.c_void_p, C.POINTER(Counts)]),
    "vsdmo_mesh_get_positions": (C.c_int, [C.c_void_p, _fp, C.c_int32]),
    "vsdmo_mesh_set_positions": (C.c_int, [C.c_void_p, _fp, C.c_int32]),
    "vsdmo_mesh_get_permutation": (C.c_int, [C.c_void_p, _ip]),
    "vsdmo_mesh_laplace": (C.c_int, [C.c_void_p, C.c_int32]),
    "vsdmo_mesh_get_cells": (C.c_int, [C.c_void_p, _ip]),
    "vsdmo_mesh_get_coloring": (C.c_int, [C.c_void_p, C.c_int32, _ip]),
    "vsdmo_mesh_get_colors": (C.c_int, [C.c_void_p, _ip]),
    "vsdmo_mesh_check_coloring": (C.c_int, [C.c_void_p, _ip, C.c_int32]),
    "vsdmo_mesh_get_faces": (C.c_int, [C.c_void_p, _ip]),
    "vsdmo_mesh_get_onering": (C.c_int, [C.c_void_p, _ip, _ip]),
    "vsdmo_mesh_get_halfedges": (C.c_int, [C.c_void_p, _ip, _ip, _ip, _ip, _ip]),
    "vsdmo_mesh_get_flags": (C.c_int, [C.c_void_p, _u8p, _u8p]),
    "vsdmo_opt_create": (C.c_int, [C.c_void_p, C.c_int32, C.c_float, C.c_int32, C.POINTER(Params),
                                   C.POINTER(C.c_void_p)]),
    "vsdmo_opt_iterate": (C.c_int, [C.c_void_p, C.c_int32]),
    "vsdmo_opt_optimize": (C.c_int, [C.c_void_p]),
    "vsdmo_opt_is_done": (C.c_int32, [C.c_void_p]),
    "vsdmo_opt_min_quality": (C.c_int, [C.c_void_p, _fp]),
    "vsdmo_opt_mean_quality": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "vsdmo_opt_histogram": (C.c_int, [C.c_void_p, C.c_int32, _ip]),
    "vsdmo_opt_histogram_noalloc": (C.c_int, [C.c_void_p, C.c_int32, _ip]),
    "vsdmo_opt_element_qualities": (C.c_int, [C.c_void_p, _fp]),
    "vsdmo_opt_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "vsdmo_opt_set_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "vsdmo_opt_set_failure_dump": (C.c_int, [C.c_void_p, C.c_char_p]),
    "vsdmo_opt_destroy": (C.c_int, [C.c_void_p]),
    "vsdmo_last_error": (C.c_char_p, []),
    "vsdmo_mesh_file_read": (C.c_int, [C.c_int, C.c_char_p, _ip, _ip, _fp, _ip]),
    "vsdmo_mesh_file_write": (C.c_int, [C.c_int, C.c_char_p, _fp, C.c_int32, _ip, C.c_int32]),
    "vsdmo_mesh_create_from_file": (C.c_int, [C.c_int, C.c_char_p, C.POINTER(MeshParams), C.POINTER(C.c_void_p)]),
    "vsdmo_mesh_write": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32]),
    "vsdmo_mlp_num_params": (C.c_int32, []),
    "vsdmo_mlp_create": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "vsdmo_mlp_destroy": (C.c_int, [C.c_void_p]),
    "vsdmo_mlp_get_params": (C.c_int, [C.c_void_p, _fp]),
    "vsdmo_mlp_set_params": (C.c_int, [C.c_void_p, _fp]),
    "vsdmo_mlp_forward": (C.c_int, [C.c_void_p, _fp, C.c_int32, _fp]),
    "vsdmo_mlp_gradients": (C.c_int, [C.c_void_p, _fp, _fp, C.c_int32, C.POINTER(C.c_double), _fp]),
    "vsdmo_mlp_train": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_uint64,
                                  C.POINTER(C.c_double)]),
    "vsdmo_dataset_create": (C.c_int, [C.c_int64, C.POINTER(C.c_void_p)]),
    "vsdmo_dataset_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "vsdmo_dataset_get": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, _fp, _fp]),
    "vsdmo_dataset_destroy": (C.c_int, [C.c_void_p]),
    "vsdmo_opt_record": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32]),
    "vsdmo_opt_surface_points": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _fp]),
    "vsdmo_opt_set_proposal": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_float]),
    "vsdmo_opt_set_fallback": (C.c_int, [C.c_void_p, C.c_int32]),
    "vsdmo_batch_params_default": (None, [C.POINTER(BatchParams)]),
    "vsdmo_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(BatchParams),
                                     C.POINTER(C.c_void_p)]),
    "vsdmo_batch_iterate": (C.c_int, [C.c_void_p, C.c_int32]),
    "vsdmo_batch_min_quality": (C.c_int, [C.c_void_p, _fp]),
    "vsdmo_batch_mean_quality": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "vsdmo_batch_get_positions": (C.c_int, [C.c_void_p, C.c_int32, _fp]),
    "vsdmo_batch_writeback": (C.c_int, [C.c_void_p]),
    "vsdmo_batch_restore": (C.c_int, [C.c_void_p]),
    "vsdmo_batch_get_params": (C.c_int, [C.c_void_p, C.POINTER(BatchParams)]),
    "vsdmo_batch_set_params": (C.c_int, [C.c_void_p, C.POINTER(BatchParams)]),
    "vsdmo_batch_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.c_float, C.c_char_p, C.POINTER(BatchParams)]),
    "vsdmo_batch_get_stats": (C.c_int, [C.c_void_p, C.POINTER(BatchStats)]),
    "vsdmo_batch_set_proposal": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_int32, C.c_float]),
    "vsdmo_batch_set_fallback": (C.c_int, [C.c_void_p, C.c_int32]),
    "vsdmo_batch_learned_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "vsdmo_batch_destroy": (C.c_int, [C.c_void_p]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Loads libvsdmo.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libvsdmo.so not found at {path}; build it with "
                           "`python -m volumesurfacediscretemeshoptimization_amd.build`")
    lib = C.CDLL(path)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class VsdmoError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str):
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def check(status: int, where: str):
    if status != VSDMO_OK:
        msg = load().vsdmo_last_error()
        raise VsdmoError(status, where, msg.decode() if msg else "")

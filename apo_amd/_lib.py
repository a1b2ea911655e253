"""ctypes loader for apo_amd/lib/libapo_hip.so (C ABI in include/apo_c.h)."""
from __future__ import annotations

import ctypes
import os
from ctypes import byref, c_uint64  # noqa: F401  (re-exported)

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libapo_hip.so")
_lib = None

_STATUS = {0: "APO_OK", 1: "APO_ERR_HIP", 2: "APO_ERR_OUT_OF_MEMORY", 3: "APO_ERR_INVALID",
           4: "APO_ERR_STATE", 5: "APO_ERR_CAPACITY"}


class ApoError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{_STATUS.get(status, status)}: {message}")
        self.status = status


class Config(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint32), ("max_iterations", ctypes.c_uint32), ("max_restarts", ctypes.c_uint32),
                ("stage_mask", ctypes.c_uint32), ("waves_per_cu", ctypes.c_uint32), ("device", ctypes.c_uint32),
                ("proposal_cost", ctypes.c_float), ("graph", ctypes.c_uint32), ("cover_ratio", ctypes.c_float),
                ("waves_cells", ctypes.c_uint32), ("waves_vertex", ctypes.c_uint32), ("waves_edge", ctypes.c_uint32),
                ("waves_bisector", ctypes.c_uint32), ("proposal_fallback_k", ctypes.c_uint32),
                ("proposal_max_k", ctypes.c_uint32), ("knn_passes", ctypes.c_uint32), ("complete", ctypes.c_uint32),
                ("proposal_radius", ctypes.c_uint32), ("cells_per_wave", ctypes.c_uint32),
                ("bisector_per_cell", ctypes.c_uint32)]


class Precision(ctypes.Structure):
    """apo_precision (include/apo_c.h)."""
    _fields_ = [(name, ctypes.c_uint64) for name in (
        "quadrisectors", "bad_quadrisectors", "uncertain_quadrisectors",
        "trisectors", "bad_trisectors", "uncertain_trisectors")]


class Stats(ctypes.Structure):
    _fields_ = [("ms_total", ctypes.c_double), ("ms_lbvh", ctypes.c_double), ("ms_knn", ctypes.c_double),
                ("ms_cells", ctypes.c_double), ("ms_validation", ctypes.c_double),
                ("ms_bisector", ctypes.c_double), ("ms_output", ctypes.c_double),
                ("sites", ctypes.c_uint64), ("vertices", ctypes.c_uint64), ("uncertified", ctypes.c_uint64),
                ("buried", ctypes.c_uint64), ("flagged_cells", ctypes.c_uint64),
                ("vertex_records", ctypes.c_uint64), ("edge_records", ctypes.c_uint64),
                ("ms_vertex_validation_kernel", ctypes.c_double), ("ms_edge_validation_kernel", ctypes.c_double),
                ("validation_launches", ctypes.c_uint64), ("validation_tasks", ctypes.c_uint64),
                ("bytes_cells", ctypes.c_uint64), ("bytes_vertex_validation", ctypes.c_uint64),
                ("bytes_edge_validation", ctypes.c_uint64), ("bytes_bisector_validation", ctypes.c_uint64),
                ("ms_proposal", ctypes.c_double), ("proposal", ctypes.c_uint64), ("cells_certified", ctypes.c_uint64),
                ("ms_completion", ctypes.c_double), ("completion_rounds", ctypes.c_uint64),
                ("completion_updated_cells", ctypes.c_uint64), ("completion_unrepairable", ctypes.c_uint64),
                ("completion_validation_tasks", ctypes.c_uint64), ("completion_rebuilt_cells", ctypes.c_uint64),
                ("reserved", ctypes.c_uint64 * 1)]


_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)

# symbol -> (restype, argtypes); the list is exactly what include/apo_c.h declares
SIGNATURES = {
    "apo_default_config": (None, [ctypes.POINTER(Config)]),
    "apo_joggle": (None, [_f64p, ctypes.c_uint64, ctypes.c_double]),
    "apo_joggle_f32": (None, [_f32p, ctypes.c_uint64, ctypes.c_double]),
    "apo_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.POINTER(Config)]),
    "apo_destroy": (None, [_vp]),
    "apo_last_error": (ctypes.c_char_p, [_vp]),
    "apo_build": (ctypes.c_int, [_vp, _f32p, _u32]),
    "apo_build_device": (ctypes.c_int, [_vp, _vp, _u32]),
    "apo_build_batched": (ctypes.c_int, [_vp, _f32p, _u32p, _u32]),
    "apo_build_batched_device": (ctypes.c_int, [_vp, _vp, _u32p, _u32]),
    "apo_batch_vertex_offsets": (ctypes.c_int, [_vp, _u64p]),
    "apo_build_device_static": (ctypes.c_int, [_vp, _vp, _u32, _vp, _vp, ctypes.c_uint64, _u64p]),
    "apo_launch_occupancy": (ctypes.c_int, [_vp, _u32p]),
    "apo_vertex_count": (ctypes.c_int, [_vp, _u64p]),
    "apo_get_vertices": (ctypes.c_int, [_vp, _f64p, _u32p, ctypes.c_uint64]),
    "apo_get_device_vertices": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "apo_closed_edge_count": (ctypes.c_int, [_vp, _u64p]),
    "apo_precision_check": (ctypes.c_int, [_vp, ctypes.POINTER(Precision)]),
    "apo_cavity_hit_ratios": (ctypes.c_int, [_vp, _u32, _f32p]),
    "apo_cavity_vertices": (ctypes.c_int, [_vp, ctypes.c_float, ctypes.c_float, _f32p, ctypes.c_uint64, _u64p]),
    "apo_precision_check_points": (ctypes.c_int, [_vp, _f64p, _u32p, ctypes.c_uint64, _u32, _u64p, _u64p]),
    "apo_get_closed_edges": (ctypes.c_int, [_vp, _f64p, _u32p, ctypes.c_uint64]),
    "apo_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(Stats)]),
    "apo_check_empty_spheres": (ctypes.c_int, [_vp, _f64p, _u32p, ctypes.c_uint64, ctypes.c_double, _u64p]),
    "apo_get_sorted_sites": (ctypes.c_int, [_vp, _f32p, _u32p]),
    "apo_get_lbvh": (ctypes.c_int, [_vp, _u64p, _i32p, _i32p, _i32p, _f32p, _u32p, _f32p]),
    "apo_get_artificial_sites": (ctypes.c_int, [_vp, _f64p]),
    "apo_get_knns": (ctypes.c_int, [_vp, _u32p]),
    "apo_get_cell_counts": (ctypes.c_int, [_vp, _u32p, _u32p, _u8p, _f32p, _u32p]),
    "apo_get_all_cells": (ctypes.c_int, [_vp, _u32p, _u32p]),
    "apo_get_topology_sets": (ctypes.c_int, [_vp, _u32p, ctypes.c_uint64, _u32p, ctypes.c_uint64, _u32p,
                                             ctypes.c_uint64, _u64p]),
    "apo_get_cell_topology": (ctypes.c_int, [_vp, _u64p, _u32p, ctypes.c_uint64, _u64p, _u32p, ctypes.c_uint64]),
    "apo_proposal_labels": (ctypes.c_int, [_vp, _f32p, _u32, _f32p, _u8p]),
    "apo_proposal_train": (ctypes.c_int, [_vp, _f32p, _u8p, _f32p, ctypes.c_uint64, _u32, _u32, ctypes.c_float,
                                          _u32, _f32p, _f32p, _f32p]),
    "apo_proposal_gradient": (ctypes.c_int, [_vp, _f32p, _f32p, _u8p, _f32p, _u32, _f32p, _f32p]),
    "apo_proposal_set": (ctypes.c_int, [_vp, _f32p]),
    "apo_proposal_infer": (ctypes.c_int, [_vp, _f32p, _u32, _u8p, _f32p]),
    "apo_proposal_classes": (ctypes.c_int, [_vp, _u8p]),
    "apo_proposal_radii": (ctypes.c_int, [_vp, _f32p]),
    "apo_proposal_set_classes": (ctypes.c_int, [_vp, _u8p, _u32]),
    "apo_proposal_set_radii": (ctypes.c_int, [_vp, _f32p, _u32]),
    "apo_cell_radii": (ctypes.c_int, [_vp, _f32p]),
}


def library_path() -> str:
    return _LIB_PATH


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise ApoError(1, f"{_LIB_PATH} is missing: build it with python -m apo_amd.build")
        L = ctypes.CDLL(_LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def create(cfg: Config):
    ctx = _vp()
    st = lib().apo_create(byref(ctx), byref(cfg))
    if st != 0:
        raise ApoError(st, "apo_create failed (no usable HIP device?)")
    return ctx


def check(ctx, status: int):
    if status != 0:
        msg = lib().apo_last_error(ctx)
        raise ApoError(status, msg.decode() if msg else "")


def fptr(a: np.ndarray):
    return a.ctypes.data_as(_f32p)


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_f64p)


def uptr(a: np.ndarray):
    return a.ctypes.data_as(_u32p)


def u64ptr(a: np.ndarray):
    return a.ctypes.data_as(_u64p)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_i32p)


def u8ptr(a: np.ndarray):
    return a.ctypes.data_as(_u8p)

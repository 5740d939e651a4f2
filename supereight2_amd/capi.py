"""ctypes binding of libse8.so (the C-ABI declared in include/se8.h).

This module is the Python check harness's view of the library: it loads the in-tree
libse8.so and raises if it is missing (there is no CPU fallback).  Device buffers are torch
tensors on the current HIP device; their data_ptr() is handed to the C-ABI.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libse8.so")

SE8_OK = 0
SE8_ERR_INVALID = -1
SE8_ERR_OOM = -2
SE8_ERR_HIP = -3
SE8_ERR_CAPACITY = -4
SE8_ERR_IO = -5
STORAGE_F32 = 0
STORAGE_F16 = 1
STORAGE_MULTIRES = 2
COPY_H2D = 0
COPY_D2H = 1
COPY_D2D = 2


TRACKER_MAX_LEVELS = 6  # SE8_TRACKER_MAX_LEVELS


class MapConfig(C.Structure):
    _fields_ = [("dim", C.c_float * 3), ("res", C.c_float), ("t_MW", C.c_float * 3)]


class TsdfConfig(C.Structure):
    _fields_ = [
        ("truncation_boundary_factor", C.c_float),
        ("max_weight", C.c_float),
        ("storage", C.c_int),
        ("max_blocks", C.c_int64),
    ]


class Pinhole(C.Structure):
    _fields_ = [
        ("width", C.c_int),
        ("height", C.c_int),
        ("fx", C.c_float),
        ("fy", C.c_float),
        ("cx", C.c_float),
        ("cy", C.c_float),
        ("near_plane", C.c_float),
        ("far_plane", C.c_float),
    ]


class LidarConfig(C.Structure):
    _fields_ = [
        ("width", C.c_int),
        ("height", C.c_int),
        ("near_plane", C.c_float),
        ("far_plane", C.c_float),
        ("beam_elevation_deg", C.POINTER(C.c_float)),
        ("beam_azimuth_deg", C.POINTER(C.c_float)),
    ]


class TrackerConfig(C.Structure):
    _fields_ = [
        ("iterations", C.c_int * TRACKER_MAX_LEVELS),
        ("num_levels", C.c_int),
        ("dist_threshold", C.c_float),
        ("normal_threshold", C.c_float),
        ("track_threshold", C.c_float),
        ("icp_threshold", C.c_float),
        ("compensated", C.c_int),
    ]


class MapInfo(C.Structure):
    _fields_ = [
        ("octree_size", C.c_int),
        ("max_scale", C.c_int),
        ("res", C.c_float),
        ("dim", C.c_float),
        ("storage", C.c_int),
        ("num_blocks", C.c_int64),
        ("num_nodes", C.c_int64),
        ("max_blocks", C.c_int64),
        ("aabb_valid", C.c_int),
        ("aabb_min", C.c_int * 3),
        ("aabb_max", C.c_int * 3),
    ]


class StageTimes(C.Structure):
    _fields_ = [
        ("fetch_frustum", C.c_float),
        ("create_list", C.c_float),
        ("allocate_list", C.c_float),
        ("update", C.c_float),
        ("raycast", C.c_float),
        ("tracking", C.c_float),
        ("last_frustum_blocks", C.c_int64),
        ("last_new_blocks", C.c_int64),
        ("last_candidates", C.c_int64),
        ("sum_fetch_frustum", C.c_double),
        ("sum_create_list", C.c_double),
        ("sum_allocate_list", C.c_double),
        ("sum_update", C.c_double),
        ("sum_raycast", C.c_double),
        ("sum_tracking", C.c_double),
        ("n_integrations", C.c_int64),
        ("n_raycasts", C.c_int64),
        ("n_tracks", C.c_int64),
        ("sum_update_blocks", C.c_int64),
        ("sum_candidates", C.c_int64),
        ("sum_track_level", C.c_double * TRACKER_MAX_LEVELS),
        ("n_track_level_launches", C.c_int64 * TRACKER_MAX_LEVELS),
    ]


TIMING_OFF, TIMING_STAGES, TIMING_LEVELS = 0, 1, 2


class ReaderConfig(C.Structure):
    _fields_ = [
        ("type", C.c_int),
        ("sequence_path", C.c_char_p),
        ("ground_truth_file", C.c_char_p),
        ("inverse_scale", C.c_float),
    ]


READER_RAW = 0
READER_TUM = 1

_P = C.c_void_p
_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int)

# name -> (restype, argtypes); every function declared in include/se8.h
SIGNATURES = {
    "se8_default_configs": (C.c_int, [C.POINTER(MapConfig), C.POINTER(TsdfConfig), C.POINTER(Pinhole), C.POINTER(TrackerConfig)]),
    "se8_map_create": (C.c_int, [C.POINTER(MapConfig), C.POINTER(TsdfConfig), C.POINTER(_P)]),
    "se8_map_destroy": (C.c_int, [_P]),
    "se8_map_get_info": (C.c_int, [_P, C.POINTER(MapInfo)]),
    "se8_integrate_depth": (C.c_int, [_P, C.c_int, _P, C.POINTER(Pinhole), _F, _P]),
    "se8_lidar_create": (C.c_int, [C.POINTER(LidarConfig), C.POINTER(_P)]),
    "se8_lidar_destroy": (C.c_int, [_P]),
    "se8_integrate_range": (C.c_int, [_P, C.c_int, _P, _P, _F, _P]),
    "se8_lidar_project": (C.c_int, [_P, _F, C.c_int64, _F, C.POINTER(C.c_uint8)]),
    "se8_raycast": (C.c_int, [_P, C.POINTER(Pinhole), _F, _P, _P, _P, _P]),
    "se8_tracker_create": (C.c_int, [C.POINTER(Pinhole), C.POINTER(TrackerConfig), C.POINTER(_P)]),
    "se8_tracker_destroy": (C.c_int, [_P]),
    "se8_track": (C.c_int, [_P, _P, _F, _P, _P, _I, _F, _P]),
    "se8_tracker_read_pyramid": (C.c_int, [_P, C.c_int, _F, _F]),
    "se8_track_raycast": (C.c_int, [_P, _P, _P, _F, _I, _P]),
    "se8_icp_cost_grad_batched": (C.c_int, [_P, _P, _P, _P, _F, C.c_int, _F, _F, _F, _F, _I, _P]),
    "se8_icp_cost_grad_batched_noalloc": (C.c_int, [_P, _P, _P, _P, _F, C.c_int, _P, _P, _P, _P, _P, _P]),
    "se8_raycast_depth_grad_batched": (C.c_int, [_P, C.POINTER(Pinhole), C.c_int, _F, _P, _P, _P, _P]),
    "se8_integrate_depth_dev": (C.c_int, [_P, C.c_int, _P, C.POINTER(Pinhole), _P, _P]),
    "se8_integrate_range_dev": (C.c_int, [_P, C.c_int, _P, _P, _P, _P]),
    "se8_raycast_dev": (C.c_int, [_P, C.POINTER(Pinhole), _P, _P, _P, _P, _P]),
    "se8_raycast_dev_prepare_track": (C.c_int, [_P, C.POINTER(Pinhole), _P, _P, _P, _P, _P, _P, _P]),
    "se8_track_dev_prepare_integrate": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "se8_track_dev": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "se8_frame_graph_create": (C.c_int, [_P, _P, C.POINTER(Pinhole), _P, _P, _P, _P, C.POINTER(_P)]),
    "se8_frame_graph_launch": (C.c_int, [_P, C.c_int, _P, _P, _P, _P, _P]),
    "se8_frame_graph_destroy": (C.c_int, [_P]),
    "se8_allocate_blocks": (C.c_int, [_P, _I, C.c_int64, C.c_int, _I, C.POINTER(C.c_int64)]),
    "se8_write_block": (C.c_int, [_P, _I, _F, _F]),
    "se8_read_block": (C.c_int, [_P, _I, _F, _F]),
    "se8_read_update_list": (C.c_int, [_P, C.POINTER(C.c_uint64), C.c_int64, C.POINTER(C.c_int64)]),
    "se8_dump_blocks": (C.c_int, [_P, C.POINTER(C.c_uint64), _I, _F, _F, C.c_int64, C.POINTER(C.c_int64)]),
    "se8_dump_blocks_multires": (C.c_int, [_P, C.POINTER(C.c_uint64), _I, _F, _F, C.c_int64, C.POINTER(C.c_int64)]),
    "se8_dump_nodes": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint8), _I, C.c_int64, C.POINTER(C.c_int64)]),
    "se8_get_data": (C.c_int, [_P, _F, C.c_int64, _F, _F]),
    "se8_field_interp": (C.c_int, [_P, _F, C.c_int64, _F, C.POINTER(C.c_uint8)]),
    "se8_field_grad": (C.c_int, [_P, _F, C.c_int64, _F, C.POINTER(C.c_uint8)]),
    "se8_mesh": (C.c_int, [_P, _F, _P, C.c_int64, C.POINTER(C.c_int64), _P]),
    "se8_save_mesh": (C.c_int, [_F, C.c_int64, C.c_int, C.POINTER(C.c_int8), C.c_char_p, _F]),
    "se8_has_supported_mesh_extension": (C.c_int, [C.c_char_p]),
    "se8_save_field_slices": (C.c_int, [_P, C.c_char_p, C.c_char_p, C.c_char_p, _F]),
    "se8_field_slice": (C.c_int, [_P, C.c_int, C.c_int, _P, _P]),
    "se8_structure": (C.c_int, [_P, _F, C.c_int, _P, _P, C.c_int64, C.POINTER(C.c_int64), _P]),
    "se8_downsample_depth": (C.c_int, [_P, C.c_int, C.c_int, _P, C.c_int, C.c_int, _P, _P]),
    "se8_downsample_depth_u16": (C.c_int, [_P, C.c_int, C.c_int, C.c_float, C.c_int, _P, C.c_int, C.c_int, _P, _P]),
    "se8_reader_create": (C.c_int, [C.POINTER(ReaderConfig), C.POINTER(_P)]),
    "se8_reader_destroy": (C.c_int, [_P]),
    "se8_reader_info": (C.c_int, [_P, _I, _I, C.POINTER(C.c_int64), _I]),
    "se8_reader_next_raw": (C.c_int, [_P, C.POINTER(C.c_uint16), _F, _I, _F, _I]),
    "se8_reader_next": (C.c_int, [_P, _F, _F, _I]),
    "se8_reader_restart": (C.c_int, [_P]),
    "se8_set_stage_timing": (C.c_int, [_P, _P, C.c_int]),
    "se8_stage_times_get": (C.c_int, [_P, _P, C.POINTER(StageTimes)]),
    "se8_stage_times_reset": (C.c_int, [_P, _P]),
    "se8_device_alloc": (C.c_int, [C.c_size_t, C.POINTER(_P)]),
    "se8_device_free": (C.c_int, [_P]),
    "se8_copy": (C.c_int, [_P, _P, C.c_size_t, C.c_int, _P]),
    "se8_stream_synchronize": (C.c_int, [_P]),
    "se8_last_error": (C.c_char_p, []),
}

_lib = None


class Se8Error(RuntimeError):
    pass


def load(path: str = LIB_PATH) -> C.CDLL:
    """Loads libse8.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise Se8Error(f"{path} is missing: build it with supereight2_amd.build.build() (no CPU fallback exists)")
    # torch ships its own HIP runtime; load it first so that libse8 binds to the same runtime
    # instance that owns the torch tensors handed to it (a second runtime in the process cannot
    # allocate on the device once the first holds it)
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != SE8_OK:
        msg = _lib.se8_last_error().decode() if _lib is not None else ""
        raise Se8Error(f"{what} failed with status {rc}: {msg}")


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_F)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_I)


def pose16(T) -> C.Array:
    T = np.asarray(T, dtype=np.float32).reshape(4, 4)
    return (C.c_float * 16)(*T.reshape(-1).tolist())


def pinhole(width=640, height=480, fx=525.0, fy=525.0, cx=319.5, cy=239.5, near_plane=0.4, far_plane=6.0) -> Pinhole:
    return Pinhole(width, height, fx, fy, cx, cy, near_plane, far_plane)

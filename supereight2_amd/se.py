"""Python mirror of the reference's C++ API for the single-res TSDF path, over libse8.so.

Names follow the reference (se::TSDFMap, se::MapIntegrator::integrateDepth,
se::raycaster::raycast_volume, se::Tracker::track, se::PinholeCamera).  Device images are torch
tensors on the current HIP device; results stay in HBM.  This is the check harness's interface;
the drop-in for C++ callers is include/se8/supereight.hpp over the same C-ABI.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import capi


def _torch():
    import torch

    return torch


def _stream_handle(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _dev_f32(t, shape=None):
    torch = _torch()
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
        raise ValueError("expected a contiguous float32 tensor on the GPU")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"expected shape {shape}, got {tuple(t.shape)}")
    return C.c_void_p(t.data_ptr())


def _dev_out(t, dtypes, numel, what):
    """Device pointer of an optional output tensor the kernel writes numel elements of: None passes
    through; otherwise a contiguous CUDA tensor of one of dtypes with exactly numel elements."""
    if t is None:
        return None
    torch = _torch()
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous()):
        raise ValueError(f"{what} must be a contiguous CUDA tensor")
    if t.dtype not in dtypes:
        raise ValueError(f"{what} must have dtype {' or '.join(str(d) for d in dtypes)}, got {t.dtype}")
    if t.numel() != numel:
        raise ValueError(f"{what} must have {numel} elements, got {t.numel()}")
    return C.c_void_p(t.data_ptr())


def _index_dtypes():
    torch = _torch()
    return (torch.int64,) + ((torch.uint64,) if hasattr(torch, "uint64") else ())


def _dev_pose(T):
    """Device pointer of a pose held in device memory (a float32 CUDA tensor of 16 values, row-major
    4x4), or None for a host pose."""
    torch = _torch()
    if isinstance(T, torch.Tensor) and T.is_cuda:
        if T.dtype != torch.float32 or T.numel() != 16 or not T.is_contiguous():
            raise ValueError("a device pose is a contiguous float32 tensor of 16 values")
        return C.c_void_p(T.data_ptr())
    return None


@dataclass
class PinholeCamera:
    """se::PinholeCamera::Config (pinhole_camera.hpp:24-40 + sensor.hpp:30-45)."""

    width: int = 640
    height: int = 480
    fx: float = 525.0
    fy: float = 525.0
    cx: float = 319.5
    cy: float = 239.5
    near_plane: float = 0.4
    far_plane: float = 6.0

    def c(self) -> capi.Pinhole:
        return capi.Pinhole(self.width, self.height, self.fx, self.fy, self.cx, self.cy, self.near_plane, self.far_plane)


class OusterLidar:
    """se::OusterLidar (ouster_lidar.hpp:12-80, ouster_lidar.cpp:59-85): a spinning lidar of
    `height` beams, each with an elevation and an azimuth offset in degrees (row 0 the highest
    beam), `width` columns over 360 degrees.  The beam model lives on the device (direction table,
    elevation and azimuth tables) from construction to close()."""

    def __init__(self, width, height, near_plane, far_plane, beam_elevation_angles, beam_azimuth_angles=None):
        self.lib = capi.load()
        self.width, self.height = int(width), int(height)
        self.near_plane, self.far_plane = float(near_plane), float(far_plane)
        self.elev = np.ascontiguousarray(beam_elevation_angles, dtype=np.float32)
        az = np.zeros(self.height) if beam_azimuth_angles is None else beam_azimuth_angles
        self.az = np.ascontiguousarray(az, dtype=np.float32)
        if self.elev.shape != (self.height,) or self.az.shape != (self.height,):
            raise ValueError("one elevation and one azimuth angle per beam are required")
        cfg = capi.LidarConfig(self.width, self.height, self.near_plane, self.far_plane, capi.fptr(self.elev),
                               capi.fptr(self.az))
        self.h = C.c_void_p()
        capi.check(self.lib.se8_lidar_create(C.byref(cfg), C.byref(self.h)), "se8_lidar_create")

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.se8_lidar_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def project(self, points_S):
        """Projects sensor-frame points (n x 3, host) with the device beam model: (uv n x 2, ok n)."""
        p = np.ascontiguousarray(np.asarray(points_S, dtype=np.float32).reshape(-1, 3))
        uv = np.zeros((len(p), 2), np.float32)
        ok = np.zeros(len(p), np.uint8)
        capi.check(self.lib.se8_lidar_project(self.h, capi.fptr(p), len(p), capi.fptr(uv),
                                              ok.ctypes.data_as(C.POINTER(C.c_uint8))), "se8_lidar_project")
        return uv, ok.astype(bool)


@dataclass
class TrackerConfig:
    """se::TrackerConfig (tracker.hpp:22-33) + the compensated-reduction switch."""

    iterations: tuple = (10, 5, 4)
    dist_threshold: float = 0.1
    normal_threshold: float = 0.8
    track_threshold: float = 0.15
    icp_threshold: float = 1e-5
    compensated: bool = False

    def c(self) -> capi.TrackerConfig:
        it = list(self.iterations)
        if not 1 <= len(it) <= capi.TRACKER_MAX_LEVELS:
            raise ValueError(f"1 to {capi.TRACKER_MAX_LEVELS} pyramid levels are supported, got {len(it)}")
        padded = it + [0] * (capi.TRACKER_MAX_LEVELS - len(it))
        return capi.TrackerConfig((C.c_int * capi.TRACKER_MAX_LEVELS)(*padded), len(it), self.dist_threshold,
                                  self.normal_threshold, self.track_threshold, self.icp_threshold,
                                  int(self.compensated))


class TSDFMap:
    """se::TSDFMap<se::Res::Single> (map.hpp:621-622): Map(dim, res, data_config)."""

    def __init__(self, dim=10.24, res=0.01, truncation_boundary_factor=8.0, max_weight=100.0, storage="f32",
                 max_blocks=0, t_MW=None):
        self.lib = capi.load()
        dims = (dim, dim, dim) if np.isscalar(dim) else tuple(dim)
        t = tuple(x / 2 for x in dims) if t_MW is None else tuple(t_MW)
        self.map_cfg = capi.MapConfig((C.c_float * 3)(*dims), res, (C.c_float * 3)(*t))
        st = {"f32": capi.STORAGE_F32, "f16": capi.STORAGE_F16, "multires": capi.STORAGE_MULTIRES}[storage]
        self.tsdf_cfg = capi.TsdfConfig(truncation_boundary_factor, max_weight, st, int(max_blocks))
        self.h = C.c_void_p()
        capi.check(self.lib.se8_map_create(C.byref(self.map_cfg), C.byref(self.tsdf_cfg), C.byref(self.h)),
                   "se8_map_create")
        self.storage = storage

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.se8_map_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> capi.MapInfo:
        out = capi.MapInfo()
        rc = self.lib.se8_map_get_info(self.h, C.byref(out))
        if rc not in (capi.SE8_OK, capi.SE8_ERR_CAPACITY):
            capi.check(rc, "se8_map_get_info")
        return out

    def getRes(self):
        return self.map_cfg.res

    def octree_size(self):
        return self.info().octree_size

    def aabb(self):
        i = self.info()
        return bool(i.aabb_valid), np.array(i.aabb_min[:]), np.array(i.aabb_max[:])

    def allocate_blocks(self, voxel_coords, only_new=False):
        c = np.ascontiguousarray(voxel_coords, dtype=np.int32).reshape(-1, 3)
        out = np.zeros((max(len(c), 1), 3), np.int32)
        n = C.c_int64()
        capi.check(self.lib.se8_allocate_blocks(self.h, capi.iptr(c), len(c), int(only_new), capi.iptr(out),
                                                C.byref(n)), "se8_allocate_blocks")
        return out[: n.value].copy()

    def write_block(self, block_coord, tsdf, weight):
        bc = np.ascontiguousarray(block_coord, dtype=np.int32)
        t = np.ascontiguousarray(tsdf, dtype=np.float32).reshape(512)
        w = np.ascontiguousarray(weight, dtype=np.float32).reshape(512)
        capi.check(self.lib.se8_write_block(self.h, capi.iptr(bc), capi.fptr(t), capi.fptr(w)), "se8_write_block")

    def read_block(self, block_coord):
        bc = np.ascontiguousarray(block_coord, dtype=np.int32)
        t = np.zeros(512, np.float32)
        w = np.zeros(512, np.float32)
        capi.check(self.lib.se8_read_block(self.h, capi.iptr(bc), capi.fptr(t), capi.fptr(w)), "se8_read_block")
        return t, w

    def update_list(self):
        """Keys of the blocks the last integration updated (integrateDepth's updated_octants), sorted."""
        n = C.c_int64()
        capi.check(self.lib.se8_read_update_list(self.h, None, 0, C.byref(n)), "se8_read_update_list")
        keys = np.zeros(max(n.value, 1), np.uint64)
        capi.check(self.lib.se8_read_update_list(self.h, keys.ctypes.data_as(C.POINTER(C.c_uint64)), len(keys),
                                                 C.byref(n)), "se8_read_update_list")
        return np.sort(keys[: n.value])

    def dump_blocks(self, with_data=True):
        n = C.c_int64()
        capi.check(self.lib.se8_dump_blocks(self.h, None, None, None, None, 0, C.byref(n)), "se8_dump_blocks")
        N = n.value
        keys = np.zeros(max(N, 1), np.uint64)
        ts = np.zeros(max(N, 1), np.int32)
        tsdf = np.zeros((max(N, 1), 512), np.float32) if with_data else None
        w = np.zeros((max(N, 1), 512), np.float32) if with_data else None
        capi.check(self.lib.se8_dump_blocks(self.h, keys.ctypes.data_as(C.POINTER(C.c_uint64)), capi.iptr(ts),
                                            capi.fptr(tsdf) if with_data else None,
                                            capi.fptr(w) if with_data else None, N, C.byref(n)), "se8_dump_blocks")
        if with_data:
            return keys[:N], ts[:N], tsdf[:N], w[:N]
        return keys[:N], ts[:N], None, None

    def dump_blocks_multires(self):
        """A multi-resolution map's blocks sorted by key: (keys, scales [N, 2] (current, minimum),
        data [N, 585, 2], past data [N, 585, 2]) with entries of scales 0..3 (block.hpp:66-166)."""
        n = C.c_int64()
        capi.check(self.lib.se8_dump_blocks_multires(self.h, None, None, None, None, 0, C.byref(n)),
                   "se8_dump_blocks_multires")
        N = n.value
        keys = np.zeros(max(N, 1), np.uint64)
        sc = np.zeros((max(N, 1), 2), np.int32)
        data = np.zeros((max(N, 1), 585, 2), np.float32)
        past = np.zeros((max(N, 1), 585, 2), np.float32)
        capi.check(self.lib.se8_dump_blocks_multires(self.h, keys.ctypes.data_as(C.POINTER(C.c_uint64)), capi.iptr(sc),
                                                     capi.fptr(data), capi.fptr(past), N, C.byref(n)),
                   "se8_dump_blocks_multires")
        return keys[:N], sc[:N], data[:N], past[:N]

    def dump_nodes(self):
        n = C.c_int64()
        capi.check(self.lib.se8_dump_nodes(self.h, None, None, None, 0, C.byref(n)), "se8_dump_nodes")
        N = n.value
        keys = np.zeros(max(N, 1), np.uint64)
        masks = np.zeros(max(N, 1), np.uint8)
        ts = np.zeros(max(N, 1), np.int32)
        capi.check(self.lib.se8_dump_nodes(self.h, keys.ctypes.data_as(C.POINTER(C.c_uint64)),
                                           masks.ctypes.data_as(C.POINTER(C.c_uint8)), capi.iptr(ts), N, C.byref(n)),
                   "se8_dump_nodes")
        return keys[:N], masks[:N], ts[:N]

    def mesh(self, T_OW=None, min_desired_scale=0, stream=None, out=None):
        """Map::mesh (map.hpp:396-404, map_impl.hpp:452-458): the surface as an (F, 3, 3) float32
        CUDA tensor of triangle vertices in frame O, metres (T_OW: host 4x4, default identity).
        out: an optional float32 CUDA tensor of shape (C, 3, 3) to write into when the mesh fits (no
        second counting pass).  min_desired_scale is unused by a single-resolution map, as in the
        reference."""
        torch = _torch()
        T = None if T_OW is None else capi.pose16(T_OW)
        st = _stream_handle(stream)
        n = C.c_int64()
        if out is not None:
            if not (out.is_cuda and out.dtype == torch.float32 and out.is_contiguous() and out.dim() == 3
                    and tuple(out.shape[1:]) == (3, 3)):
                raise ValueError("out must be a contiguous float32 CUDA tensor of shape (C, 3, 3)")
            rc = self.lib.se8_mesh(self.h, T, C.c_void_p(out.data_ptr()), out.shape[0], C.byref(n), st)
            if rc == capi.SE8_OK:
                return out[: n.value]
            if rc != capi.SE8_ERR_CAPACITY:
                capi.check(rc, "se8_mesh")
        else:
            capi.check(self.lib.se8_mesh(self.h, T, None, 0, C.byref(n), st), "se8_mesh")
        F = n.value
        faces = torch.empty((max(F, 1), 3, 3), dtype=torch.float32, device="cuda")
        if F:
            capi.check(self.lib.se8_mesh(self.h, T, C.c_void_p(faces.data_ptr()), F, C.byref(n), st), "se8_mesh")
        return faces[:F]

    def saveMesh(self, filename, T_OW=None, min_desired_scale=0):
        """Map::saveMesh (map.hpp:406-413, map_impl.hpp:462-469): writes mesh(T_OW) as .ply, .vtk or
        .obj by the extension; returns 0 on success and 1 if the file cannot be written, and raises
        ValueError for another extension (the reference throws std::invalid_argument)."""
        torch = _torch()
        faces = self.mesh(T_OW, min_desired_scale)
        torch.cuda.current_stream().synchronize()
        return save_mesh(faces.cpu().numpy(), filename)

    def saveFieldSlices(self, filename_x, filename_y, filename_z, point_W):
        """Map::saveFieldSlices (map.hpp:309-325, map_impl.hpp:272-308): VTK planes of the tsdf through
        the voxel of point_W; an empty filename skips that plane.  Returns 0 (SE8_ERR_IO raises)."""
        p = np.ascontiguousarray(point_W, dtype=np.float32).reshape(3)
        capi.check(self.lib.se8_save_field_slices(self.h, str(filename_x or "").encode(), str(filename_y or "").encode(),
                                                  str(filename_z or "").encode(), capi.fptr(p)), "se8_save_field_slices")
        return 0

    def field_slice(self, axis, index, stream=None):
        """One plane of saveFieldSlices on the device: [size, size] float32 CUDA tensor, rows = the slower
        free coordinate."""
        torch = _torch()
        n = self.octree_size()
        out = torch.empty((n, n), dtype=torch.float32, device="cuda")
        capi.check(self.lib.se8_field_slice(self.h, int(axis), int(index), C.c_void_p(out.data_ptr()),
                                            _stream_handle(stream)), "se8_field_slice")
        return out

    def structure(self, T_OW=None, only_leaves=True, stream=None):
        """Map::structure (map.hpp:415-421, map_impl.hpp:473-479): (quads [Q, 4, 3] float32 CUDA tensor,
        scales [Q] int8 CUDA tensor), six quads per octant (only the leaves by default, as the reference)."""
        torch = _torch()
        T = None if T_OW is None else capi.pose16(T_OW)
        st = _stream_handle(stream)
        n = C.c_int64()
        capi.check(self.lib.se8_structure(self.h, T, int(only_leaves), None, None, 0, C.byref(n), st), "se8_structure")
        Q = n.value
        quads = torch.empty((max(Q, 1), 4, 3), dtype=torch.float32, device="cuda")
        scales = torch.empty(max(Q, 1), dtype=torch.int8, device="cuda")
        if Q:
            capi.check(self.lib.se8_structure(self.h, T, int(only_leaves), C.c_void_p(quads.data_ptr()),
                                              C.c_void_p(scales.data_ptr()), Q, C.byref(n), st), "se8_structure")
        return quads[:Q], scales[:Q]

    def saveStructure(self, filename, T_OW=None, only_leaves=True):
        """Map::saveStructure (map.hpp:422-430, map_impl.hpp:484-490)."""
        torch = _torch()
        q, sc = self.structure(T_OW, only_leaves)
        torch.cuda.current_stream().synchronize()
        return save_mesh(q.cpu().numpy(), filename, scales=sc.cpu().numpy())

    def getData(self, points_W):
        p = np.ascontiguousarray(points_W, dtype=np.float32).reshape(-1, 3)
        t = np.zeros(len(p), np.float32)
        w = np.zeros(len(p), np.float32)
        capi.check(self.lib.se8_get_data(self.h, capi.fptr(p), len(p), capi.fptr(t), capi.fptr(w)), "se8_get_data")
        return t, w

    def getFieldInterp(self, points_W):
        p = np.ascontiguousarray(points_W, dtype=np.float32).reshape(-1, 3)
        out = np.zeros(len(p), np.float32)
        valid = np.zeros(len(p), np.uint8)
        capi.check(self.lib.se8_field_interp(self.h, capi.fptr(p), len(p), capi.fptr(out),
                                             valid.ctypes.data_as(C.POINTER(C.c_uint8))), "se8_field_interp")
        return out, valid.astype(bool)

    def getFieldGrad(self, points_W):
        p = np.ascontiguousarray(points_W, dtype=np.float32).reshape(-1, 3)
        out = np.zeros((len(p), 3), np.float32)
        valid = np.zeros(len(p), np.uint8)
        capi.check(self.lib.se8_field_grad(self.h, capi.fptr(p), len(p), capi.fptr(out),
                                           valid.ctypes.data_as(C.POINTER(C.c_uint8))), "se8_field_grad")
        return out, valid.astype(bool)


def has_supported_mesh_extension(filename: str) -> bool:
    """se::io::has_supported_mesh_extension (mesh_io.hpp:24-29)."""
    return bool(capi.load().se8_has_supported_mesh_extension(str(filename).encode()))


def save_mesh(faces, filename, T_OM=None, scales=None):
    """se::io::save_mesh (mesh_io.hpp:44-46, impl/mesh_io_impl.hpp:14-230) of host faces (F, 3, 3)
    triangles or (F, 4, 3) quads, with optional per-face scales (F int8): 0 on success, 1 if the
    file cannot be written; ValueError for an unsupported extension."""
    lib = capi.load()
    fa = np.ascontiguousarray(faces, dtype=np.float32)
    nv = fa.shape[1] if fa.ndim == 3 else 3
    f = fa.reshape(-1, 3 * nv)
    sc = None if scales is None else np.ascontiguousarray(scales, dtype=np.int8)
    T = None if T_OM is None else capi.pose16(T_OM)
    rc = lib.se8_save_mesh(capi.fptr(f) if len(f) else None, len(f), nv,
                           sc.ctypes.data_as(C.POINTER(C.c_int8)) if sc is not None else None,
                           str(filename).encode(), T)
    if rc == capi.SE8_ERR_IO:
        return 1
    if rc == capi.SE8_ERR_INVALID:
        raise ValueError(lib.se8_last_error().decode())
    capi.check(rc, "se8_save_mesh")
    return 0


def downsample_depth(input_depth, output_depth, downsample_map=None, stream=None):
    """se::preprocessor::downsample_depth (preprocessor.cpp:19-84) on device tensors: input [H, W]
    float32, output [H/r, W/r] float32; downsample_map (optional, [H/r, W/r] int64/uint64 CUDA
    tensor) receives the index map the reference returns."""
    h_in, w_in = input_depth.shape
    h_out, w_out = output_depth.shape
    mp = _dev_out(downsample_map, _index_dtypes(), h_out * w_out, "downsample_map")
    capi.check(capi.load().se8_downsample_depth(_dev_f32(input_depth), w_in, h_in, _dev_f32(output_depth), w_out,
                                                h_out, mp, _stream_handle(stream)), "se8_downsample_depth")
    return downsample_map


def downsample_depth_u16(raw_depth, scale, divide, output_depth, downsample_map=None, stream=None):
    """downsample_depth over depth as a reader stores it (uint16 [H, W] CUDA tensor), converted to
    metres as the reader would (raw / scale if divide, else raw * scale) inside the same kernel."""
    torch = _torch()
    if not (raw_depth.is_cuda and raw_depth.dtype in (torch.uint16, torch.int16) and raw_depth.is_contiguous()):
        raise ValueError("raw_depth must be a contiguous 16-bit CUDA tensor")
    h_in, w_in = raw_depth.shape
    h_out, w_out = output_depth.shape
    mp = _dev_out(downsample_map, _index_dtypes(), h_out * w_out, "downsample_map")
    capi.check(capi.load().se8_downsample_depth_u16(C.c_void_p(raw_depth.data_ptr()), w_in, h_in, float(scale),
                                                    int(divide), _dev_f32(output_depth), w_out, h_out, mp,
                                                    _stream_handle(stream)), "se8_downsample_depth_u16")
    return downsample_map


class ReaderStatus:
    """se::ReaderStatus (app/include/reader_base.hpp:45-62)."""

    ok, skip, eof, error = 0, 1, 2, 3


class Reader:
    """se::RAWReader / se::TUMReader (app/include/reader_raw.hpp, reader_tum.hpp) over se8_reader_*:
    reader_type "raw" or "tum".  Construction raises capi.Se8Error where the reference's reader
    would start in the error status."""

    def __init__(self, reader_type, sequence_path, ground_truth_file="", inverse_scale=0.0):
        self.lib = capi.load()
        t = {"raw": capi.READER_RAW, "tum": capi.READER_TUM}[reader_type]
        self._cfg = capi.ReaderConfig(t, str(sequence_path).encode(), str(ground_truth_file or "").encode(),
                                      float(inverse_scale))
        self.h = C.c_void_p()
        capi.check(self.lib.se8_reader_create(C.byref(self._cfg), C.byref(self.h)), "se8_reader_create")
        w, hh, gt = C.c_int(), C.c_int(), C.c_int()
        n = C.c_int64()
        capi.check(self.lib.se8_reader_info(self.h, C.byref(w), C.byref(hh), C.byref(n), C.byref(gt)),
                   "se8_reader_info")
        self.width, self.height, self.num_frames, self.has_ground_truth = w.value, hh.value, n.value, bool(gt.value)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.se8_reader_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nextData(self, with_pose=True):
        """Reader::nextData: (status, depth [H, W] float32 metres, T_WB 4x4 or None)."""
        depth = np.zeros((self.height, self.width), np.float32)
        T = np.zeros(16, np.float32)
        st = C.c_int()
        capi.check(self.lib.se8_reader_next(self.h, capi.fptr(depth), capi.fptr(T) if with_pose else None,
                                            C.byref(st)), "se8_reader_next")
        return st.value, depth, (T.reshape(4, 4) if with_pose and st.value in (0, 1) else None)

    def nextDataRaw(self, with_pose=True):
        """The next frame as stored: (status, depth [H, W] uint16, scale, divide, T_WB or None)."""
        depth = np.zeros((self.height, self.width), np.uint16)
        T = np.zeros(16, np.float32)
        scale, divide, st = C.c_float(), C.c_int(), C.c_int()
        capi.check(self.lib.se8_reader_next_raw(self.h, depth.ctypes.data_as(C.POINTER(C.c_uint16)), C.byref(scale),
                                                C.byref(divide), capi.fptr(T) if with_pose else None, C.byref(st)),
                   "se8_reader_next_raw")
        return st.value, depth, scale.value, divide.value, (T.reshape(4, 4) if with_pose and st.value in (0, 1) else None)

    def restart(self):
        capi.check(self.lib.se8_reader_restart(self.h), "se8_reader_restart")


class MapIntegrator:
    """se::MapIntegrator<MapT> (map_integrator.hpp:66-81)."""

    def __init__(self, map_: TSDFMap):
        self.map = map_

    def integrateDepth(self, timestamp, depth, sensor, T_WC, stream=None):
        """depth: [H, W] float32 device tensor: z-depth for a PinholeCamera, range for an OusterLidar.
        T_WC: a host 4x4 pose, or a device pose (float32 CUDA tensor of 16 values) that the
        integration reads on the GPU."""
        d = _dev_f32(depth, (sensor.height, sensor.width))
        Td = _dev_pose(T_WC)
        lib, st = self.map.lib, _stream_handle(stream)
        if isinstance(sensor, OusterLidar):
            if Td is not None:
                capi.check(lib.se8_integrate_range_dev(self.map.h, int(timestamp), d, sensor.h, Td, st),
                           "se8_integrate_range_dev")
            else:
                capi.check(lib.se8_integrate_range(self.map.h, int(timestamp), d, sensor.h, capi.pose16(T_WC), st),
                           "se8_integrate_range")
            return
        if Td is not None:
            capi.check(lib.se8_integrate_depth_dev(self.map.h, int(timestamp), d, C.byref(sensor.c()), Td, st),
                       "se8_integrate_depth_dev")
        else:
            capi.check(lib.se8_integrate_depth(self.map.h, int(timestamp), d, C.byref(sensor.c()), capi.pose16(T_WC),
                                               st), "se8_integrate_depth")


def raycast_volume(map_: TSDFMap, sensor: PinholeCamera, T_WS, points_W, normals_W, scale=None, stream=None):
    """se::raycaster::raycast_volume (raycaster.hpp:56-64) into device tensors [H, W, 3]."""
    shape = (sensor.height, sensor.width, 3)
    sp = _dev_out(scale, (_torch().int8,), sensor.height * sensor.width, "scale")
    Td = _dev_pose(T_WS)
    if Td is not None:
        capi.check(map_.lib.se8_raycast_dev(map_.h, C.byref(sensor.c()), Td, _dev_f32(points_W, shape),
                                            _dev_f32(normals_W, shape), sp, _stream_handle(stream)), "se8_raycast_dev")
        return
    capi.check(map_.lib.se8_raycast(map_.h, C.byref(sensor.c()), capi.pose16(T_WS), _dev_f32(points_W, shape),
                                    _dev_f32(normals_W, shape), sp, _stream_handle(stream)), "se8_raycast")


def raycast_volume_prepare_track(map_: TSDFMap, sensor: PinholeCamera, T_WS, points_W, normals_W, tracker,
                                 next_depth, scale=None, stream=None):
    """raycast_volume at a device pose, with the pyramid of the tracker's next call (next_depth, from
    the same pose) in the same launch (se8_raycast_dev_prepare_track): the next track_device with
    that depth tensor and T_in == T_WS skips its pyramid."""
    shape = (sensor.height, sensor.width, 3)
    sp = _dev_out(scale, (_torch().int8,), sensor.height * sensor.width, "scale")
    Td = _dev_pose(T_WS)
    if Td is None:
        raise ValueError("raycast_volume_prepare_track takes a device pose")
    capi.check(map_.lib.se8_raycast_dev_prepare_track(
        map_.h, C.byref(sensor.c()), Td, _dev_f32(points_W, shape), _dev_f32(normals_W, shape), sp, tracker.h,
        _dev_f32(next_depth, (tracker.sensor.height, tracker.sensor.width)), _stream_handle(stream)),
        "se8_raycast_dev_prepare_track")


class Tracker:
    """se::Tracker<MapT, SensorT> (tracker.hpp:35-84)."""

    def __init__(self, map_: TSDFMap, sensor: PinholeCamera, config: TrackerConfig | None = None):
        self.map = map_
        self.sensor = sensor
        self.config = config or TrackerConfig()
        self.lib = capi.load()
        self.h = C.c_void_p()
        capi.check(self.lib.se8_tracker_create(C.byref(sensor.c()), C.byref(self.config.c()), C.byref(self.h)),
                   "se8_tracker_create")

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.se8_tracker_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def track(self, depth, T_WS, points_W=None, normals_W=None, stream=None):
        """Returns (ok, T_WS_new, sums32).  Without reference maps the map is raycast at T_WS."""
        T = capi.pose16(T_WS)
        ok = C.c_int()
        sums = np.zeros(32, np.float32)
        shape = (self.sensor.height, self.sensor.width)
        if points_W is None:
            capi.check(self.lib.se8_track_raycast(self.h, self.map.h, _dev_f32(depth, shape), T, C.byref(ok),
                                                  _stream_handle(stream)), "se8_track_raycast")
        else:
            capi.check(self.lib.se8_track(self.h, _dev_f32(depth, shape), T, _dev_f32(points_W, shape + (3,)),
                                          _dev_f32(normals_W, shape + (3,)), C.byref(ok), capi.fptr(sums),
                                          _stream_handle(stream)), "se8_track")
        return bool(ok.value), np.array(T[:], np.float32).reshape(4, 4), sums

    def read_pyramid(self, level):
        """The input vertex and normal maps of pyramid level `level` from the last track call
        ((H >> level, W >> level, 3) float32 each, host)."""
        h, w = self.sensor.height >> level, self.sensor.width >> level
        pts = np.zeros((h, w, 3), np.float32)
        nrm = np.zeros((h, w, 3), np.float32)
        capi.check(self.lib.se8_tracker_read_pyramid(self.h, int(level), capi.fptr(pts), capi.fptr(nrm)),
                   "se8_tracker_read_pyramid")
        return pts, nrm

    def track_device(self, depth, T_in, T_out, points_W, normals_W, ok=None, stream=None):
        """Tracker::track with device poses (float32 CUDA tensors of 16 values; T_in may be T_out):
        enqueued on the stream, nothing is read back.  ok: optional int32 CUDA tensor of one value."""
        shape = (self.sensor.height, self.sensor.width)
        Ti, To = _dev_pose(T_in), _dev_pose(T_out)
        if Ti is None or To is None:
            raise ValueError("track_device takes device poses")
        okp = _dev_out(ok, (_torch().int32,), 1, "ok")
        capi.check(self.lib.se8_track_dev(self.h, _dev_f32(depth, shape), Ti, To, _dev_f32(points_W, shape + (3,)),
                                          _dev_f32(normals_W, shape + (3,)), okp, _stream_handle(stream)),
                   "se8_track_dev")


def track_device_prepare_integrate(tracker, depth, T_in, T_out, points_W, normals_W, map_, ok=None, stream=None):
    """Tracker.track_device whose last launch also starts map_'s next integration at the tracked pose
    (se8_track_dev_prepare_integrate): the next integrateDepth on map_ with the pose tensor T_out
    skips that launch."""
    shape = (tracker.sensor.height, tracker.sensor.width)
    Ti, To = _dev_pose(T_in), _dev_pose(T_out)
    if Ti is None or To is None:
        raise ValueError("track_device_prepare_integrate takes device poses")
    okp = _dev_out(ok, (_torch().int32,), 1, "ok")
    capi.check(tracker.lib.se8_track_dev_prepare_integrate(
        tracker.h, _dev_f32(depth, shape), Ti, To, _dev_f32(points_W, shape + (3,)), _dev_f32(normals_W, shape + (3,)),
        okp, map_.h, _stream_handle(stream)), "se8_track_dev_prepare_integrate")


class FrameGraph:
    """One frame of the tracking loop (Tracker::track, integrateDepth, raycast_volume) as a replayed
    HIP graph (se8_frame_graph_*).  points/normals: two (H, W, 3) float32 CUDA tensors each; launch i
    tracks against pair (i + 1) % 2 and raycasts into pair i % 2.  Stage timing must be off."""

    def __init__(self, tracker: Tracker, map_: TSDFMap, sensor: PinholeCamera, points, normals):
        shape = (sensor.height, sensor.width, 3)
        self.lib = capi.load()
        self.h = C.c_void_p()
        self._keep = (points, normals)
        capi.check(self.lib.se8_frame_graph_create(tracker.h, map_.h, C.byref(sensor.c()),
                                                   _dev_f32(points[0], shape), _dev_f32(normals[0], shape),
                                                   _dev_f32(points[1], shape), _dev_f32(normals[1], shape),
                                                   C.byref(self.h)), "se8_frame_graph_create")
        self.shape = (sensor.height, sensor.width)

    def launch(self, timestamp, depth, T_in, T_out, ok=None, stream=None):
        Ti, To = _dev_pose(T_in), _dev_pose(T_out)
        if Ti is None or To is None:
            raise ValueError("a frame graph takes device poses")
        okp = _dev_out(ok, (_torch().int32,), 1, "ok")
        capi.check(self.lib.se8_frame_graph_launch(self.h, int(timestamp), _dev_f32(depth, self.shape), Ti, To, okp,
                                                   _stream_handle(stream)), "se8_frame_graph_launch")

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.se8_frame_graph_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stage_times(map_: TSDFMap | None = None, tracker: Tracker | None = None) -> dict:
    out = capi.StageTimes()
    lib = capi.load()
    capi.check(lib.se8_stage_times_get(map_.h if map_ else None, tracker.h if tracker else None, C.byref(out)),
               "se8_stage_times_get")
    res = {}
    for k, _ in capi.StageTimes._fields_:
        v = getattr(out, k)
        res[k] = list(v) if isinstance(v, C.Array) else v
    return res


def set_stage_timing(map_: TSDFMap | None = None, tracker: Tracker | None = None, mode: str = "stages") -> None:
    """mode "off" (no events, no host waits: the host runs ahead), "stages" (default) or "levels"
    (stages plus per-pyramid-level ICP time)."""
    m = {"off": capi.TIMING_OFF, "stages": capi.TIMING_STAGES, "levels": capi.TIMING_LEVELS}[mode]
    capi.check(capi.load().se8_set_stage_timing(map_.h if map_ else None, tracker.h if tracker else None, m),
               "se8_set_stage_timing")


def stage_times_reset(map_: TSDFMap | None = None, tracker: Tracker | None = None) -> None:
    lib = capi.load()
    capi.check(lib.se8_stage_times_reset(map_.h if map_ else None, tracker.h if tracker else None),
               "se8_stage_times_reset")

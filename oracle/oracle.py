"""ctypes loader of the CPU oracle (oracle/se8_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, to check results.  The product library never uses it.

Two builds exist: precision="f32" follows the fp32 operation order of the GPU path (parity),
precision="f64" is the FP64 twin used to report precision errors.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


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


MAX_LEVELS = 6  # OR_MAX_LEVELS


class TrackerConfig(C.Structure):
    _fields_ = [
        ("iterations", C.c_int * MAX_LEVELS),
        ("num_levels", C.c_int),
        ("dist_threshold", C.c_float),
        ("normal_threshold", C.c_float),
        ("track_threshold", C.c_float),
        ("icp_threshold", C.c_float),
    ]


def pinhole(width=640, height=480, fx=525.0, fy=525.0, cx=319.5, cy=239.5, near_plane=0.4, far_plane=6.0) -> Pinhole:
    return Pinhole(width, height, fx, fy, cx, cy, near_plane, far_plane)


def tracker_config(iterations=(10, 5, 4), dist=0.1, normal=0.8, track=0.15, icp=1e-5) -> TrackerConfig:
    it = list(iterations)
    assert 1 <= len(it) <= MAX_LEVELS
    return TrackerConfig((C.c_int * MAX_LEVELS)(*(it + [0] * (MAX_LEVELS - len(it)))), len(it), dist, normal, track,
                         icp)


_libs: dict = {}


def lib(precision: str = "f32") -> C.CDLL:
    if precision in _libs:
        return _libs[precision]
    path = os.path.join(BUILD, f"liboracle_{precision}.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    R = C.POINTER(C.c_double if precision == "f64" else C.c_float)
    P = C.c_void_p
    U64 = C.c_uint64
    I = C.POINTER(C.c_int)
    L.or_expand.restype = U64
    L.or_expand.argtypes = [U64]
    L.or_compact.restype = U64
    L.or_compact.argtypes = [U64]
    L.or_encode_code.restype = U64
    L.or_encode_code.argtypes = [C.c_int, C.c_int, C.c_int]
    L.or_decode_code.argtypes = [U64, I]
    L.or_encode_key.restype = C.c_int
    L.or_encode_key.argtypes = [I, C.c_int, C.POINTER(U64)]
    L.or_decode_key.argtypes = [U64, I, I]
    L.or_code_to_child_idx.restype = C.c_int
    L.or_code_to_child_idx.argtypes = [U64, C.c_int]
    L.or_key_at_scale.restype = C.c_int
    L.or_key_at_scale.argtypes = [U64, C.c_int, C.POINTER(U64)]
    L.or_parent_key.restype = U64
    L.or_parent_key.argtypes = [U64]
    L.or_block_key.restype = U64
    L.or_block_key.argtypes = [U64, C.c_int]
    L.or_is_child.restype = C.c_int
    L.or_is_child.argtypes = [U64, U64]
    L.or_is_siblings.restype = C.c_int
    L.or_is_siblings.argtypes = [U64, U64]
    for fn in ("or_unique_keys", "or_unique_codes", "or_unique_at_scale", "or_unique_allocation"):
        getattr(L, fn).restype = C.c_long
    L.or_unique_keys.argtypes = [C.POINTER(U64), C.c_long, C.POINTER(U64)]
    L.or_unique_codes.argtypes = [C.POINTER(U64), C.c_long, C.POINTER(U64)]
    L.or_unique_at_scale.argtypes = [C.POINTER(U64), C.c_long, C.c_int, C.POINTER(U64)]
    L.or_unique_allocation.argtypes = [C.POINTER(U64), C.c_long, C.c_int, C.POINTER(U64)]
    L.or_map_create.restype = P
    L.or_map_create.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float]
    L.or_map_destroy.argtypes = [P]
    L.or_map_size.restype = C.c_int
    L.or_map_size.argtypes = [P]
    L.or_map_num_blocks.restype = C.c_long
    L.or_map_num_blocks.argtypes = [P]
    L.or_map_num_nodes.restype = C.c_long
    L.or_map_num_nodes.argtypes = [P]
    L.or_map_dump_blocks.argtypes = [P, C.POINTER(U64), I, R, R]
    L.or_map_last_update.restype = C.c_long
    L.or_map_last_update.argtypes = [P, C.POINTER(U64)]
    L.or_map_dump_nodes.restype = C.c_long
    L.or_map_dump_nodes.argtypes = [P, C.POINTER(U64), C.POINTER(C.c_uint8), I]
    L.or_map_aabb.restype = C.c_int
    L.or_map_aabb.argtypes = [P, I]
    L.or_allocate_blocks.restype = C.c_long
    L.or_allocate_blocks.argtypes = [P, I, C.c_long, C.c_int, I]
    L.or_set_block_data.restype = C.c_int
    L.or_set_block_data.argtypes = [P, I, R, R]
    L.or_set_multires_block.restype = C.c_int
    L.or_set_multires_block.argtypes = [P, I, C.c_int, C.c_int, R, R]
    L.or_mesh.restype = C.c_long
    L.or_mesh.argtypes = [P, R, R, C.c_long]
    L.or_field_slice.argtypes = [P, C.c_int, C.c_int, R]
    L.or_map_set_multires.argtypes = [P]
    L.or_map_set_storage_f16.argtypes = [P]
    L.or_map_touch_begin.argtypes = [P]
    L.or_map_touch_end.argtypes = [P, C.POINTER(C.c_long), C.POINTER(U64)]
    L.or_half_round.argtypes = [C.c_float]
    L.or_half_round.restype = C.c_float
    L.or_map_dump_multires.argtypes = [P, C.POINTER(U64), I, R, R]
    L.or_structure.restype = C.c_long
    L.or_structure.argtypes = [P, R, C.c_int, R, C.POINTER(C.c_int8), C.c_long]
    L.or_depth_from_u16.argtypes = [C.POINTER(C.c_uint16), C.c_long, C.c_float, C.c_int, C.POINTER(C.c_float)]
    L.or_downsample_depth.restype = C.c_int
    L.or_downsample_depth.argtypes = [C.POINTER(C.c_float), C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int,
                                      C.c_int, C.POINTER(U64)]
    L.or_get_field_interp.restype = C.c_int
    L.or_get_field_interp.argtypes = [P, R, R]
    L.or_get_field_grad.restype = C.c_int
    L.or_get_field_grad.argtypes = [P, R, R]
    L.or_get_data.restype = C.c_int
    L.or_get_data.argtypes = [P, R, R, R]
    L.or_integrate.argtypes = [P, C.c_int, C.POINTER(C.c_float), C.POINTER(Pinhole), R, C.POINTER(C.c_long)]
    L.or_raycast.argtypes = [P, C.POINTER(Pinhole), R, R, R, C.POINTER(C.c_int8)]
    L.or_raycast_stats.argtypes = [P, C.POINTER(Pinhole), R, R, R, I]
    FP = C.POINTER(C.c_float)
    L.or_lidar_create.restype = P
    L.or_lidar_create.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, FP, FP]
    L.or_lidar_destroy.argtypes = [P]
    L.or_lidar_project.restype = C.c_int
    L.or_lidar_project.argtypes = [P, R, R]
    L.or_lidar_project_n.argtypes = [P, R, C.c_long, R, C.POINTER(C.c_uint8)]
    L.or_lidar_backproject.argtypes = [P, C.c_int, C.c_int, R]
    L.or_lidar_sphere_in_frustum.restype = C.c_int
    L.or_lidar_sphere_in_frustum.argtypes = [P, R, C.c_double if precision == "f64" else C.c_float]
    L.or_atan2_det.restype = C.c_double if precision == "f64" else C.c_float
    L.or_atan2_det.argtypes = [C.c_double if precision == "f64" else C.c_float] * 2
    L.or_sin_det.restype = C.c_double if precision == "f64" else C.c_float
    L.or_sin_det.argtypes = [C.c_double if precision == "f64" else C.c_float]
    L.or_integrate_lidar.argtypes = [P, C.c_int, FP, P, R, C.POINTER(C.c_long)]
    L.or_compute_T_CV.argtypes = [P, R, R]
    L.or_half_sample.argtypes = [C.POINTER(C.c_float), C.c_int, C.c_int, C.POINTER(C.c_float), C.c_float]
    L.or_depth_to_points.argtypes = [C.POINTER(C.c_float), C.POINTER(Pinhole), C.c_int, R]
    L.or_points_to_normals.argtypes = [R, C.c_int, C.c_int, C.c_int, R]
    L.or_track.restype = C.c_int
    L.or_track.argtypes = [C.POINTER(C.c_float), C.POINTER(Pinhole), C.POINTER(TrackerConfig), R, R, R, R, I, I]
    L.or_icp_track_reduce.argtypes = [R, R, C.c_int, C.c_int, R, R, C.POINTER(Pinhole), R, R, C.c_float,
                                      C.c_float, I, R, R, R]
    L.or_icp_solve.argtypes = [R, R]
    L.or_se3_exp.argtypes = [R, R]
    L.or_first_block.argtypes = [P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float, C.c_float,
                                 C.POINTER(C.c_float)]
    L.or_raycast_depth_grad.argtypes = [P, C.POINTER(Pinhole), R, C.c_int, R, R, C.POINTER(C.c_uint8)]
    L.or_real_size.restype = C.c_int
    L.or_set_threads.argtypes = [C.c_int]
    _libs[precision] = L
    return L


def _dtype(precision):
    return np.float64 if precision == "f64" else np.float32


def _rp(a, precision):
    return a.ctypes.data_as(C.POINTER(C.c_double if precision == "f64" else C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _u64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def set_threads(n: int) -> None:
    for p in ("f32", "f64"):
        lib(p).or_set_threads(n)


# ----------------------------------------------------------------------------------------------
# key operations
# ----------------------------------------------------------------------------------------------
def encode_key(coord, scale):
    L = lib()
    c = np.ascontiguousarray(coord, dtype=np.int32)
    k = C.c_uint64()
    full = L.or_encode_key(_ip(c), int(scale), C.byref(k))
    return int(k.value), bool(full)


def decode_key(key):
    L = lib()
    c = np.zeros(3, np.int32)
    s = C.c_int()
    L.or_decode_key(C.c_uint64(key), _ip(c), C.byref(s))
    return c, int(s.value)


def encode_code(coord):
    return int(lib().or_encode_code(int(coord[0]), int(coord[1]), int(coord[2])))


def decode_code(code):
    c = np.zeros(3, np.int32)
    lib().or_decode_code(C.c_uint64(code), _ip(c))
    return c


def _unique(fn, keys, *extra):
    k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    out = np.zeros(max(len(k), 1), np.uint64)
    n = fn(_u64p(k), len(k), *extra, _u64p(out)) if not extra else None
    return k, out, n


def unique_keys(keys):
    k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    out = np.zeros(max(len(k), 1), np.uint64)
    n = lib().or_unique_keys(_u64p(k), len(k), _u64p(out))
    return out[:n]


def unique_codes(keys):
    k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    out = np.zeros(max(len(k), 1), np.uint64)
    n = lib().or_unique_codes(_u64p(k), len(k), _u64p(out))
    return out[:n]


def unique_at_scale(keys, scale):
    k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    out = np.zeros(max(len(k), 1), np.uint64)
    n = lib().or_unique_at_scale(_u64p(k), len(k), int(scale), _u64p(out))
    return out[:n]


def unique_allocation(keys, max_block_scale):
    k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    out = np.zeros(max(len(k), 1), np.uint64)
    n = lib().or_unique_allocation(_u64p(k), len(k), int(max_block_scale), _u64p(out))
    return out[:n]


# ----------------------------------------------------------------------------------------------
# map
# ----------------------------------------------------------------------------------------------
class OracleMap:
    def __init__(self, dim=10.24, res=0.01, mu_factor=8.0, max_weight=100.0, precision="f32", multires=False,
                 storage="f32"):
        """storage="f16": tsdf values are rounded to the nearest binary16 after every update, as an
        FP16-TSDF map stores them (weights are integers, exact either way)."""
        self.precision = precision
        self.L = lib(precision)
        self.dtype = _dtype(precision)
        self.h = self.L.or_map_create(float(dim), float(res), float(mu_factor), float(max_weight))
        self.res = res
        self.multires = multires
        if multires:
            self.L.or_map_set_multires(self.h)
        if storage == "f16":
            self.L.or_map_set_storage_f16(self.h)
        elif storage != "f32":
            raise ValueError(f"unknown storage {storage!r}")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.or_map_destroy(self.h)
            self.h = None

    @property
    def size(self):
        return self.L.or_map_size(self.h)

    @property
    def num_blocks(self):
        return self.L.or_map_num_blocks(self.h)

    @property
    def num_nodes(self):
        return self.L.or_map_num_nodes(self.h)

    def integrate(self, ts, depth, pin, T_WC):
        d = np.ascontiguousarray(depth, dtype=np.float32)
        T = np.ascontiguousarray(np.asarray(T_WC).reshape(16), dtype=self.dtype)
        stats = (C.c_long * 4)()
        self.L.or_integrate(self.h, int(ts), d.ctypes.data_as(C.POINTER(C.c_float)), C.byref(pin),
                            _rp(T, self.precision), stats)
        return dict(frustum=stats[0], new=stats[1], candidates=stats[2], valid_px=stats[3])

    def integrate_lidar(self, ts, rng, lidar, T_WS):
        """integrateDepth with an OusterLidar; rng is the H x W range image (metres)."""
        r = np.ascontiguousarray(rng, dtype=np.float32)
        assert r.shape == (lidar.height, lidar.width)
        T = np.ascontiguousarray(np.asarray(T_WS).reshape(16), dtype=self.dtype)
        stats = (C.c_long * 4)()
        self.L.or_integrate_lidar(self.h, int(ts), r.ctypes.data_as(C.POINTER(C.c_float)), lidar.handle(self.precision),
                                  _rp(T, self.precision), stats)
        return dict(frustum=stats[0], new=stats[1], candidates=stats[2], valid_px=stats[3])

    def raycast_touched(self, pin, T_WS, bits=False):
        """raycast() that also counts the distinct voxels whose data it reads and the distinct blocks
        holding them: (points, normals, scale, voxels, blocks), and with bits the read mask (8 uint64
        words per block in allocation order, bit i of a block = its voxel i)."""
        n_blocks = self.num_blocks
        self.L.or_map_touch_begin(self.h)
        p, n, s = self.raycast(pin, T_WS)
        out = (C.c_long * 2)()
        mask = np.zeros(max(n_blocks, 1) * 8, np.uint64)
        self.L.or_map_touch_end(self.h, out, _u64p(mask) if bits else None)
        if bits:
            return p, n, s, int(out[0]), int(out[1]), mask[: n_blocks * 8].reshape(n_blocks, 8)
        return p, n, s, int(out[0]), int(out[1])

    def raycast_stats(self, pin, T_WS):
        """Per-pixel march statistics (H, W, 7): traversal iterations, samples in absent bricks, in
        unallocated blocks of present bricks, at unobserved and at observed voxels, interpolations, and
        of the samples in unallocated blocks of present bricks those whose 32-voxel octant is absent."""
        T = np.ascontiguousarray(np.asarray(T_WS).reshape(16), dtype=self.dtype)
        pts = np.zeros((pin.height, pin.width, 3), self.dtype)
        nrm = np.zeros_like(pts)
        st = np.zeros((pin.height, pin.width, 7), np.int32)
        self.L.or_raycast_stats(self.h, C.byref(pin), _rp(T, self.precision), _rp(pts, self.precision),
                                _rp(nrm, self.precision), st.ctypes.data_as(C.POINTER(C.c_int)))
        return st

    def raycast(self, pin, T_WS):
        T = np.ascontiguousarray(np.asarray(T_WS).reshape(16), dtype=self.dtype)
        pts = np.zeros((pin.height, pin.width, 3), self.dtype)
        nrm = np.zeros((pin.height, pin.width, 3), self.dtype)
        scale = np.full((pin.height, pin.width), -1, np.int8)
        self.L.or_raycast(self.h, C.byref(pin), _rp(T, self.precision), _rp(pts, self.precision),
                          _rp(nrm, self.precision), scale.ctypes.data_as(C.POINTER(C.c_int8)))
        return pts, nrm, scale

    def dump_blocks(self, with_data=True):
        n = self.num_blocks
        keys = np.zeros(max(n, 1), np.uint64)
        ts = np.zeros(max(n, 1), np.int32)
        tsdf = np.zeros((max(n, 1), 512), self.dtype) if with_data else None
        w = np.zeros((max(n, 1), 512), self.dtype) if with_data else None
        R = C.POINTER(C.c_double if self.precision == "f64" else C.c_float)
        self.L.or_map_dump_blocks(self.h, _u64p(keys), _ip(ts), _rp(tsdf, self.precision) if with_data else R(),
                                  _rp(w, self.precision) if with_data else R())
        if with_data:
            return keys[:n], ts[:n], tsdf[:n], w[:n]
        return keys[:n], ts[:n], None, None

    def update_list(self):
        """Keys of the blocks the last integration updated (allocated + frustum: integrateDepth's
        updated_octants, map_integrator_impl.hpp:77-80), sorted."""
        n = self.L.or_map_last_update(self.h, C.POINTER(C.c_uint64)())
        keys = np.zeros(max(n, 1), np.uint64)
        self.L.or_map_last_update(self.h, _u64p(keys))
        return np.sort(keys[:n])

    def dump_multires(self):
        """(keys, scales [N, 2] (current, minimum), data [N, 585, 2], past data [N, 585, 2]) sorted by key."""
        n = self.num_blocks
        keys = np.zeros(max(n, 1), np.uint64)
        sc = np.zeros((max(n, 1), 2), np.int32)
        data = np.zeros((max(n, 1), 585, 2), self.dtype)
        past = np.zeros((max(n, 1), 585, 2), self.dtype)
        self.L.or_map_dump_multires(self.h, _u64p(keys), _ip(sc), _rp(data, self.precision), _rp(past, self.precision))
        return keys[:n], sc[:n], data[:n], past[:n]

    def dump_nodes(self):
        n = self.num_nodes
        keys = np.zeros(max(n, 1), np.uint64)
        masks = np.zeros(max(n, 1), np.uint8)
        ts = np.zeros(max(n, 1), np.int32)
        k = self.L.or_map_dump_nodes(self.h, _u64p(keys), masks.ctypes.data_as(C.POINTER(C.c_uint8)), _ip(ts))
        order = np.argsort(keys[:k], kind="stable")
        return keys[:k][order], masks[:k][order], ts[:k][order]

    def aabb(self):
        out = np.zeros(6, np.int32)
        valid = self.L.or_map_aabb(self.h, _ip(out))
        return bool(valid), out[:3].copy(), out[3:].copy()

    def allocate_blocks(self, voxel_coords, only_new=False):
        c = np.ascontiguousarray(voxel_coords, dtype=np.int32).reshape(-1, 3)
        out = np.zeros((max(len(c), 1), 3), np.int32)
        n = self.L.or_allocate_blocks(self.h, _ip(c), len(c), int(only_new), _ip(out))
        return out[:n].copy()

    def set_block_data(self, block_coord, tsdf, weight):
        bc = np.ascontiguousarray(block_coord, dtype=np.int32)
        t = np.ascontiguousarray(tsdf, dtype=self.dtype)
        w = np.ascontiguousarray(weight, dtype=self.dtype)
        return self.L.or_set_block_data(self.h, _ip(bc), _rp(t, self.precision), _rp(w, self.precision))

    def set_multires_block(self, block_coord, cur_scale, min_scale, data, past):
        """Sets an allocated block's multi-resolution state: data and past data [585, 2] (tsdf,
        weight) in dump_multires's layout."""
        bc = np.ascontiguousarray(block_coord, dtype=np.int32)
        d = np.ascontiguousarray(data, dtype=self.dtype).reshape(585, 2)
        p = np.ascontiguousarray(past, dtype=self.dtype).reshape(585, 2)
        return self.L.or_set_multires_block(self.h, _ip(bc), int(cur_scale), int(min_scale), _rp(d, self.precision),
                                            _rp(p, self.precision))

    def load_blocks(self, block_coords, tsdf, weight):
        """Allocates the given blocks and sets their 512 voxels each (e.g. a dump of another map)."""
        bc = np.ascontiguousarray(block_coords, dtype=np.int32).reshape(-1, 3)
        if len(bc):
            self.allocate_blocks(bc)
        for i in range(len(bc)):
            assert self.set_block_data(bc[i], tsdf[i], weight[i]) == 0

    def mesh(self, T_OW=None):
        """Map::mesh (map_impl.hpp:452-458): (F, 3, 3) triangle vertices in frame O, in the
        reference's loop order over this map's blocks."""
        T = np.eye(4, dtype=self.dtype) if T_OW is None else np.ascontiguousarray(T_OW, dtype=self.dtype)
        R = C.POINTER(C.c_double if self.precision == "f64" else C.c_float)
        n = self.L.or_mesh(self.h, _rp(T, self.precision), R(), 0)
        faces = np.zeros((max(n, 1), 3, 3), self.dtype)
        n2 = self.L.or_mesh(self.h, _rp(T, self.precision), _rp(faces, self.precision), n)
        assert n2 == n
        return faces[:n]

    def field_slice(self, axis, index):
        """One plane of Map::saveFieldSlices: [size, size] tsdf values (1 where unallocated)."""
        out = np.zeros((self.size, self.size), self.dtype)
        self.L.or_field_slice(self.h, int(axis), int(index), _rp(out, self.precision))
        return out

    def structure(self, T_OW=None, only_leaves=True):
        """Map::structure: (quads [Q, 4, 3], scales [Q]) in the octree's depth-first order."""
        T = np.eye(4, dtype=self.dtype) if T_OW is None else np.ascontiguousarray(T_OW, dtype=self.dtype)
        R = C.POINTER(C.c_double if self.precision == "f64" else C.c_float)
        n = self.L.or_structure(self.h, _rp(T, self.precision), int(only_leaves), R(), C.POINTER(C.c_int8)(), 0)
        q = np.zeros((max(n, 1), 4, 3), self.dtype)
        sc = np.zeros(max(n, 1), np.int8)
        n2 = self.L.or_structure(self.h, _rp(T, self.precision), int(only_leaves), _rp(q, self.precision),
                                 sc.ctypes.data_as(C.POINTER(C.c_int8)), n)
        assert n2 == n
        return q[:n], sc[:n]

    def field_interp(self, p):
        pp = np.ascontiguousarray(p, dtype=self.dtype)
        out = np.zeros(1, self.dtype)
        ok = self.L.or_get_field_interp(self.h, _rp(pp, self.precision), _rp(out, self.precision))
        return (float(out[0]) if ok else None)

    def field_grad(self, p):
        pp = np.ascontiguousarray(p, dtype=self.dtype)
        out = np.zeros(3, self.dtype)
        ok = self.L.or_get_field_grad(self.h, _rp(pp, self.precision), _rp(out, self.precision))
        return (out.copy() if ok else None)

    def get_data(self, p):
        pp = np.ascontiguousarray(p, dtype=self.dtype)
        t = np.zeros(1, self.dtype)
        w = np.zeros(1, self.dtype)
        self.L.or_get_data(self.h, _rp(pp, self.precision), _rp(t, self.precision), _rp(w, self.precision))
        return float(t[0]), float(w[0])

    def first_block(self, o, d, near_t, far_t):
        oo = np.ascontiguousarray(o, dtype=np.float32)
        dd = np.ascontiguousarray(d, dtype=np.float32)
        out = np.zeros(2, np.float32)
        fp = C.POINTER(C.c_float)
        self.L.or_first_block(self.h, oo.ctypes.data_as(fp), dd.ctypes.data_as(fp), float(near_t), float(far_t),
                              out.ctypes.data_as(fp))
        return float(out[0]), float(out[1])

    def raycast_depth_grad(self, pin, T_WS, iters=40):
        """Depth of the refined field root per pixel and d depth / d xi (left twist [v; w])."""
        T = np.ascontiguousarray(np.asarray(T_WS).reshape(16), dtype=self.dtype)
        depth = np.zeros((pin.height, pin.width), self.dtype)
        grad = np.zeros((pin.height, pin.width, 6), self.dtype)
        valid = np.zeros((pin.height, pin.width), np.uint8)
        self.L.or_raycast_depth_grad(self.h, C.byref(pin), _rp(T, self.precision), int(iters),
                                     _rp(depth, self.precision), _rp(grad, self.precision),
                                     valid.ctypes.data_as(C.POINTER(C.c_uint8)))
        return depth, grad, valid.astype(bool)

    def T_CV(self, T_WC):
        T = np.ascontiguousarray(np.asarray(T_WC).reshape(16), dtype=self.dtype)
        out = np.zeros(16, self.dtype)
        self.L.or_compute_T_CV(self.h, _rp(T, self.precision), _rp(out, self.precision))
        return out.reshape(4, 4)


def track(depth, pin, T_WS, ref_pts, ref_nrm, cfg=None, precision="f32", with_level_iterations=False):
    L = lib(precision)
    dt = _dtype(precision)
    cfg = cfg or tracker_config()
    d = np.ascontiguousarray(depth, dtype=np.float32)
    T = np.ascontiguousarray(np.asarray(T_WS).reshape(16), dtype=dt).copy()
    rp = np.ascontiguousarray(ref_pts, dtype=dt)
    rn = np.ascontiguousarray(ref_nrm, dtype=dt)
    sums = np.zeros(32, dt)
    iters = C.c_int()
    per_level = (C.c_int * MAX_LEVELS)()
    ok = L.or_track(d.ctypes.data_as(C.POINTER(C.c_float)), C.byref(pin), C.byref(cfg), _rp(T, precision),
                    _rp(rp, precision), _rp(rn, precision), _rp(sums, precision), C.byref(iters), per_level)
    if with_level_iterations:
        return bool(ok), T.reshape(4, 4), sums, int(iters.value), list(per_level)[: cfg.num_levels]
    return bool(ok), T.reshape(4, 4), sums, int(iters.value)


def half_round(x):
    """The nearest binary16 value of each float32 in x (ties to even), as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    out = np.array([lib().or_half_round(float(v)) for v in a.reshape(-1)], np.float32)
    return out.reshape(a.shape)


def depth_from_u16(raw, scale, divide):
    """Readers' conversion of stored uint16 depth: raw / scale (RAW, millimetres, scale 1000) or
    raw * scale (TUM inverse scale), in float32."""
    r = np.ascontiguousarray(raw, dtype=np.uint16)
    out = np.zeros(r.shape, np.float32)
    lib().or_depth_from_u16(r.ctypes.data_as(C.POINTER(C.c_uint16)), r.size, float(scale), int(divide),
                            out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def downsample_depth(depth, w_out, h_out):
    """preprocessor::downsample_depth: (output depth [h_out, w_out], index map [h_out, w_out] uint64)."""
    d = np.ascontiguousarray(depth, dtype=np.float32)
    h_in, w_in = d.shape
    out = np.zeros((h_out, w_out), np.float32)
    mp = np.zeros((h_out, w_out), np.uint64)
    rc = lib().or_downsample_depth(d.ctypes.data_as(C.POINTER(C.c_float)), w_in, h_in,
                                   out.ctypes.data_as(C.POINTER(C.c_float)), w_out, h_out, _u64p(mp))
    if rc != 0:
        raise ValueError("downsample_depth: the input must be the same integer multiple of the output")
    return out, mp


def half_sample(depth, e_d=0.1 * 3):
    d = np.ascontiguousarray(depth, dtype=np.float32)
    h, w = d.shape
    out = np.zeros((h // 2, w // 2), np.float32)
    lib().or_half_sample(d.ctypes.data_as(C.POINTER(C.c_float)), w, h, out.ctypes.data_as(C.POINTER(C.c_float)),
                         float(np.float32(0.1) * np.float32(3)))
    return out


def depth_to_points(depth, pin, level, precision="f32"):
    dt = _dtype(precision)
    d = np.ascontiguousarray(depth, dtype=np.float32)
    h, w = d.shape
    out = np.zeros((h, w, 3), dt)
    lib(precision).or_depth_to_points(d.ctypes.data_as(C.POINTER(C.c_float)), C.byref(pin), int(level),
                                      _rp(out, precision))
    return out


def points_to_normals(pts, neg_y=False, precision="f32"):
    dt = _dtype(precision)
    p = np.ascontiguousarray(pts, dtype=dt)
    h, w, _ = p.shape
    out = np.zeros_like(p)
    lib(precision).or_points_to_normals(_rp(p, precision), w, h, int(neg_y), _rp(out, precision))
    return out


def icp_track_reduce(in_pts, in_nrm, ref_pts, ref_nrm, pin, T_WS, T_ref, dist=0.1, normal=0.8, precision="f32"):
    dt = _dtype(precision)
    ip = np.ascontiguousarray(in_pts, dtype=dt)
    inn = np.ascontiguousarray(in_nrm, dtype=dt)
    h, w, _ = ip.shape
    rp = np.ascontiguousarray(ref_pts, dtype=dt)
    rn = np.ascontiguousarray(ref_nrm, dtype=dt)
    T = np.ascontiguousarray(np.asarray(T_WS).reshape(16), dtype=dt)
    Tr = np.ascontiguousarray(np.asarray(T_ref).reshape(16), dtype=dt)
    res = np.zeros((h, w), np.int32)
    err = np.zeros((h, w), dt)
    J = np.zeros((h, w, 6), dt)
    sums = np.zeros(32, dt)
    lib(precision).or_icp_track_reduce(_rp(ip, precision), _rp(inn, precision), w, h, _rp(rp, precision),
                                       _rp(rn, precision), C.byref(pin), _rp(T, precision), _rp(Tr, precision),
                                       float(dist), float(normal), _ip(res), _rp(err, precision), _rp(J, precision),
                                       _rp(sums, precision))
    return res, err, J, sums


def icp_solve(vals27, precision="f32"):
    dt = _dtype(precision)
    v = np.ascontiguousarray(vals27, dtype=dt)
    x = np.zeros(6, dt)
    lib(precision).or_icp_solve(_rp(v, precision), _rp(x, precision))
    return x


def se3_exp(a, precision="f32"):
    dt = _dtype(precision)
    v = np.ascontiguousarray(a, dtype=dt)
    T = np.zeros(16, dt)
    lib(precision).or_se3_exp(_rp(v, precision), _rp(T, precision))
    return T.reshape(4, 4)


class OracleLidar:
    """OusterLidar (ouster_lidar.cpp:59-100) with the projection model stated in se8_oracle.c: beam
    elevations strictly descending (row 0 highest), azimuth offsets per beam, in degrees."""

    def __init__(self, width, height, near_plane, far_plane, elevation_deg, azimuth_deg=None):
        self.width, self.height = int(width), int(height)
        self.near_plane, self.far_plane = float(near_plane), float(far_plane)
        self.elev = np.ascontiguousarray(elevation_deg, dtype=np.float32)
        self.az = np.ascontiguousarray(np.zeros(height) if azimuth_deg is None else azimuth_deg, dtype=np.float32)
        assert self.elev.shape == (height,) and self.az.shape == (height,)
        self._h = {}

    def handle(self, precision="f32"):
        if precision not in self._h:
            L = lib(precision)
            FP = C.POINTER(C.c_float)
            h = L.or_lidar_create(self.width, self.height, self.near_plane, self.far_plane, self.elev.ctypes.data_as(FP),
                                  self.az.ctypes.data_as(FP))
            if not h:
                raise ValueError("lidar config rejected (need height >= 2 and strictly descending elevations)")
            self._h[precision] = h
        return self._h[precision]

    def __del__(self):
        for prec, h in getattr(self, "_h", {}).items():
            lib(prec).or_lidar_destroy(h)

    def project(self, pts, precision="f32"):
        dt = _dtype(precision)
        p = np.ascontiguousarray(pts, dtype=dt).reshape(-1, 3)
        uv = np.zeros((len(p), 2), dt)
        ok = np.zeros(len(p), np.uint8)
        lib(precision).or_lidar_project_n(self.handle(precision), _rp(p, precision), len(p), _rp(uv, precision),
                                          ok.ctypes.data_as(C.POINTER(C.c_uint8)))
        return uv, ok.astype(bool)

    def backproject(self, u, v, precision="f32"):
        out = np.zeros(3, _dtype(precision))
        lib(precision).or_lidar_backproject(self.handle(precision), int(u), int(v), _rp(out, precision))
        return out

    def directions(self, precision="f32"):
        """All back-projection directions, [H, W, 3]."""
        out = np.zeros((self.height, self.width, 3), _dtype(precision))
        for v in range(self.height):
            for u in range(self.width):
                out[v, u] = self.backproject(u, v, precision)
        return out

    def sphere_in_frustum(self, c, r, precision="f32"):
        cc = np.ascontiguousarray(c, dtype=_dtype(precision))
        return bool(lib(precision).or_lidar_sphere_in_frustum(self.handle(precision), _rp(cc, precision), float(r)))


def atan2_det(y, x, precision="f32"):
    return lib(precision).or_atan2_det(float(y), float(x))


def sin_det(x, precision="f32"):
    return lib(precision).or_sin_det(float(x))

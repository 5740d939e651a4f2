"""Seeded synthetic depth workloads for the single-resolution TSDF pipeline.

Scene: an axis-aligned box room x in [-3,3], y in [-2,2], z in [0,3] (world frame, z up) with
three spheres of radius U(0.3, 0.8) placed inside it, clear of the camera path.  Depth is
ray-cast analytically for a pinhole camera (640x480, fx=fy=525, cx=319.5, cy=239.5) whose frame
is z forward, x right, y down; depth noise is N(0, (0.0012 + 0.0019 (z - 0.4)^2)^2) and 1% of
the pixels are set to 0.

Trajectories:
  circle_trajectory  - the mapping-only sequence: a circle of radius 1.5 m at 1.4 m height,
                       looking at the room centre, evenly over 360 degrees;
  spline_trajectory  - the tracking sequence: a seeded smooth path around the room centre with
                       at most 2 cm and 1 degree of motion between consecutive frames.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

ROOM_MIN = np.array([-3.0, -2.0, 0.0])
ROOM_MAX = np.array([3.0, 2.0, 3.0])
ROOM_CENTRE = (ROOM_MIN + ROOM_MAX) / 2
PATH_RADIUS = 1.5
PATH_HEIGHT = 1.4


@dataclass
class Camera:
    width: int = 640
    height: int = 480
    fx: float = 525.0
    fy: float = 525.0
    cx: float = 319.5
    cy: float = 239.5
    near_plane: float = 0.4
    far_plane: float = 6.0


@dataclass
class Scene:
    centres: np.ndarray  # (k, 3)
    radii: np.ndarray    # (k,)


def make_scene(seed: int = 0, n_spheres: int = 3) -> Scene:
    rng = np.random.default_rng(seed)
    centres, radii = [], []
    while len(centres) < n_spheres:
        r = rng.uniform(0.3, 0.8)
        c = rng.uniform(ROOM_MIN + r, ROOM_MAX - r)
        # keep clear of the camera path (a ring of radius PATH_RADIUS at PATH_HEIGHT) by 0.35 m
        d_ring = np.hypot(np.hypot(c[0], c[1]) - PATH_RADIUS, c[2] - PATH_HEIGHT)
        if d_ring < r + 0.35:
            continue
        if any(np.linalg.norm(c - c2) < r + r2 + 0.05 for c2, r2 in zip(centres, radii)):
            continue
        centres.append(c)
        radii.append(r)
    return Scene(np.array(centres), np.array(radii))


def look_at(position, target, up=(0.0, 0.0, 1.0)) -> np.ndarray:
    """T_WC (float64 4x4) of a z-forward, x-right, y-down camera at position looking at target."""
    p = np.asarray(position, np.float64)
    f = np.asarray(target, np.float64) - p
    f /= np.linalg.norm(f)
    x = np.cross(f, np.asarray(up, np.float64))
    x /= np.linalg.norm(x)
    y = np.cross(f, x)
    T = np.eye(4)
    T[:3, 0], T[:3, 1], T[:3, 2], T[:3, 3] = x, y, f, p
    return T


def circle_trajectory(n: int = 50, radius: float = PATH_RADIUS, height: float = PATH_HEIGHT) -> list:
    target = np.array([0.0, 0.0, ROOM_CENTRE[2]])
    out = []
    for k in range(n):
        a = 2.0 * np.pi * k / n
        out.append(look_at([radius * np.cos(a), radius * np.sin(a), height], target))
    return out


def _catmull_rom(ctrl: np.ndarray, t: np.ndarray) -> np.ndarray:
    """Uniform Catmull-Rom spline through ctrl (m, d) at parameters t in [0, m-1]."""
    m = len(ctrl)
    P = np.concatenate([ctrl[:1], ctrl, ctrl[-1:]], axis=0)
    i = np.clip(np.floor(t).astype(int), 0, m - 2)
    u = (t - i)[:, None]
    p0, p1, p2, p3 = P[i], P[i + 1], P[i + 2], P[i + 3]
    return 0.5 * ((2 * p1) + (-p0 + p2) * u + (2 * p0 - 5 * p1 + 4 * p2 - p3) * u**2 + (-p0 + 3 * p1 - 3 * p2 + p3) * u**3)


def rotation_angle_deg(R: np.ndarray) -> float:
    """Angle of the rotation R in degrees, as atan2 of its antisymmetric part's norm and its trace:
    accurate near the identity, where the arccos of the trace alone is lost to rounding (a float32
    rotation matrix compared with its own float64 value reads up to 0.02 degrees that way)."""
    R = np.asarray(R, dtype=np.float64)
    s = 0.5 * np.linalg.norm([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    c = (np.trace(R) - 1.0) / 2.0
    return float(np.degrees(np.arctan2(s, c)))


def spline_trajectory(n: int = 300, seed: int = 1, max_step_m: float = 0.02, max_step_deg: float = 1.0) -> list:
    """Seeded smooth path: an arc of the camera ring with seeded radius/height/gaze wobble.
    The path is slowed down uniformly until every step is within the motion limits."""
    rng = np.random.default_rng(seed)
    n_ctrl = 9
    ctrl = np.stack([
        np.linspace(0.0, 1.0, n_ctrl),                         # arc parameter
        rng.uniform(-0.15, 0.15, n_ctrl),                      # radius offset (m)
        rng.uniform(-0.10, 0.10, n_ctrl),                      # height offset (m)
        rng.uniform(-0.25, 0.25, n_ctrl),                      # gaze target x offset
        rng.uniform(-0.25, 0.25, n_ctrl),                      # gaze target y offset
        rng.uniform(-0.15, 0.15, n_ctrl),                      # gaze target z offset
    ], axis=1)
    arc = np.radians(200.0)
    a0 = rng.uniform(0, 2 * np.pi)
    speed = 1.0
    for _ in range(20):
        t = np.linspace(0.0, (n_ctrl - 1) * speed, n)
        s = _catmull_rom(ctrl, t)
        poses = []
        for row in s:
            a = a0 + arc * row[0]
            r = PATH_RADIUS + row[1]
            pos = [r * np.cos(a), r * np.sin(a), PATH_HEIGHT + row[2]]
            tgt = np.array([0.0, 0.0, ROOM_CENTRE[2]]) + row[3:6]
            poses.append(look_at(pos, tgt))
        steps = [np.linalg.norm(poses[k + 1][:3, 3] - poses[k][:3, 3]) for k in range(n - 1)]
        angs = [rotation_angle_deg(poses[k][:3, :3].T @ poses[k + 1][:3, :3]) for k in range(n - 1)]
        worst = max(max(steps) / max_step_m, max(angs) / max_step_deg)
        if worst <= 1.0:
            return poses
        speed /= worst * 1.02
    raise RuntimeError("could not satisfy the motion limits")


def render_depth(scene: Scene, cam: Camera, T_WC: np.ndarray, rng=None, noise: bool = True,
                 dropout: float = 0.01) -> np.ndarray:
    """Analytic z-depth (float32, HxW) of the room and spheres seen from T_WC."""
    u, v = np.meshgrid(np.arange(cam.width, dtype=np.float64), np.arange(cam.height, dtype=np.float64))
    dC = np.stack([(u - cam.cx) / cam.fx, (v - cam.cy) / cam.fy, np.ones_like(u)], axis=-1)
    R, o = T_WC[:3, :3], T_WC[:3, 3]
    d = dC @ R.T  # world direction, z_C component = 1 so the ray parameter is the z-depth
    with np.errstate(divide="ignore", invalid="ignore"):
        t_hi = np.where(d > 0, (ROOM_MAX - o) / d, np.where(d < 0, (ROOM_MIN - o) / d, np.inf))
    t = np.min(t_hi, axis=-1)
    for c, r in zip(scene.centres, scene.radii):
        oc = o - c
        a = np.sum(d * d, axis=-1)
        b = 2.0 * (d @ oc)
        cc = oc @ oc - r * r
        disc = b * b - 4 * a * cc
        hit = disc > 0
        sq = np.sqrt(np.where(hit, disc, 0.0))
        t0 = (-b - sq) / (2 * a)
        ok = hit & (t0 > 1e-6)
        t = np.where(ok & (t0 < t), t0, t)
    z = t
    if noise:
        rng = rng if rng is not None else np.random.default_rng(0)
        sigma = 0.0012 + 0.0019 * (z - 0.4) ** 2
        z = z + rng.normal(0.0, 1.0, z.shape) * sigma
        if dropout > 0:
            z = np.where(rng.random(z.shape) < dropout, 0.0, z)
    return z.astype(np.float32)


def sequence(kind: str = "circle", n: int = 50, seed: int = 0, cam: Camera | None = None, noise: bool = True):
    """(depth frames [n,H,W] float32, ground-truth poses [n,4,4] float64) of a seeded workload."""
    cam = cam or Camera()
    scene = make_scene(seed)
    poses = circle_trajectory(n) if kind == "circle" else spline_trajectory(n, seed=seed + 1)
    rng = np.random.default_rng(seed + 1000)
    depth = np.stack([render_depth(scene, cam, T, rng, noise=noise) for T in poses])
    return depth, np.stack(poses)


def render_sequence_torch(scene: Scene, cam: Camera, poses, seed: int = 0, noise: bool = True,
                          dropout: float = 0.01, device: str = "cuda"):
    """The same analytic depth as render_depth, computed in float64 on the GPU with torch, for
    long sequences; returns a [n, H, W] float32 device tensor.  Noise comes from a seeded torch
    generator on the device (not bit-identical to the numpy noise of render_depth)."""
    import torch

    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    f64 = torch.float64
    v, u = torch.meshgrid(torch.arange(cam.height, dtype=f64, device=device),
                          torch.arange(cam.width, dtype=f64, device=device), indexing="ij")
    dC = torch.stack([(u - cam.cx) / cam.fx, (v - cam.cy) / cam.fy, torch.ones_like(u)], dim=-1)
    lo = torch.tensor(ROOM_MIN, dtype=f64, device=device)
    hi = torch.tensor(ROOM_MAX, dtype=f64, device=device)
    centres = torch.tensor(scene.centres, dtype=f64, device=device)
    radii = scene.radii.tolist()
    out = torch.empty((len(poses), cam.height, cam.width), dtype=torch.float32, device=device)
    for k, T in enumerate(poses):
        T = torch.tensor(np.asarray(T, np.float64), dtype=f64, device=device)
        R, o = T[:3, :3], T[:3, 3]
        d = dC @ R.T
        pos_t = torch.where(d > 0, (hi - o) / d, torch.full_like(d, float("inf")))
        neg_t = torch.where(d < 0, (lo - o) / d, torch.full_like(d, float("inf")))
        t = torch.minimum(pos_t, neg_t).amin(dim=-1)
        a = (d * d).sum(-1)
        for ci in range(len(radii)):
            oc = o - centres[ci]
            b = 2.0 * (d @ oc)
            cc = oc @ oc - radii[ci] ** 2
            disc = b * b - 4 * a * cc
            hit = disc > 0
            sq = torch.sqrt(torch.clamp(disc, min=0.0))
            t0 = (-b - sq) / (2 * a)
            t = torch.where(hit & (t0 > 1e-6) & (t0 < t), t0, t)
        z = t
        if noise:
            sigma = 0.0012 + 0.0019 * (z - 0.4) ** 2
            z = z + torch.randn(z.shape, generator=gen, dtype=f64, device=device) * sigma
            if dropout > 0:
                z = torch.where(torch.rand(z.shape, generator=gen, dtype=f64, device=device) < dropout,
                                torch.zeros_like(z), z)
        out[k] = z.to(torch.float32)
    return out


# ------------------------------------------------------------------------------------------------
# Street workload for a spinning lidar: a ground plane z = 0, axis-aligned box buildings either
# side of a straight street along x, and vertical cylinders (poles, trunks) on the pavements.  The
# sensor frame is x forward, y left, z up; scans are taken at STREET_SENSOR_HEIGHT along y = 0.
# ------------------------------------------------------------------------------------------------
STREET_SENSOR_HEIGHT = 2.0


@dataclass
class LidarSpec:
    """Beam layout of the scanner: `height` beams evenly over +-fov/2 of elevation (row 0 the
    highest), no azimuth offsets, `width` columns over 360 degrees."""

    width: int = 1024
    height: int = 128
    fov_deg: float = 45.0
    near_plane: float = 0.5
    far_plane: float = 60.0

    def elevations_deg(self) -> np.ndarray:
        return np.linspace(self.fov_deg / 2, -self.fov_deg / 2, self.height).astype(np.float32)

    def azimuths_deg(self) -> np.ndarray:
        return np.zeros(self.height, np.float32)


@dataclass
class Street:
    box_lo: np.ndarray    # (k, 3)
    box_hi: np.ndarray    # (k, 3)
    cyl_xy: np.ndarray    # (m, 2)
    cyl_r: np.ndarray     # (m,)
    cyl_h: np.ndarray     # (m,)


def make_street(seed: int = 0, n_boxes: int = 40, n_cylinders: int = 200, half_length: float = 100.0) -> Street:
    """40 buildings with sides U(5, 20) m whose street-facing wall lies U(8, 14) m from the path,
    and 200 cylinders of radius U(0.1, 0.4) m and height U(3, 8) m at U(2.5, 7) m from the path."""
    rng = np.random.default_rng(seed)
    size = rng.uniform(5.0, 20.0, (n_boxes, 3))
    cx = rng.uniform(-half_length + 10.0, half_length - 10.0, n_boxes)
    side = rng.choice([-1.0, 1.0], n_boxes)
    near = rng.uniform(8.0, 14.0, n_boxes)
    cy = side * (near + size[:, 1] / 2)
    lo = np.stack([cx - size[:, 0] / 2, cy - size[:, 1] / 2, np.zeros(n_boxes)], 1)
    hi = np.stack([cx + size[:, 0] / 2, cy + size[:, 1] / 2, size[:, 2]], 1)
    px = rng.uniform(-half_length, half_length, n_cylinders)
    py = rng.choice([-1.0, 1.0], n_cylinders) * rng.uniform(2.5, 7.0, n_cylinders)
    r = rng.uniform(0.1, 0.4, n_cylinders)
    h = rng.uniform(3.0, 8.0, n_cylinders)
    return Street(lo, hi, np.stack([px, py], 1), r, h)


def street_trajectory(n: int = 200, spacing: float = 0.5) -> np.ndarray:
    """n sensor poses T_WS (float64) along y = 0 at STREET_SENSOR_HEIGHT, centred on x = 0."""
    poses = np.tile(np.eye(4), (n, 1, 1))
    poses[:, 0, 3] = (np.arange(n) - n / 2) * spacing
    poses[:, 2, 3] = STREET_SENSOR_HEIGHT
    return poses


def lidar_directions(spec: LidarSpec) -> np.ndarray:
    """Unit beam directions (H, W, 3) in float64: azimuth 2 pi u / W - pi, elevation per row."""
    e = np.radians(spec.elevations_deg().astype(np.float64))[:, None]
    phi = 2.0 * np.pi * np.arange(spec.width, dtype=np.float64)[None, :] / spec.width - np.pi
    phi = phi + np.radians(spec.azimuths_deg().astype(np.float64))[:, None]
    return np.stack([np.cos(e) * np.cos(phi), np.cos(e) * np.sin(phi), np.broadcast_to(np.sin(e), phi.shape)], -1)


def render_street_ranges(street: Street, spec: LidarSpec, poses, seed: int = 0, noise: float = 0.02,
                         device: str = "cpu"):
    """Analytic range images [n, H, W] (float32 torch tensor on `device`) of the street seen from
    each pose, with N(0, noise) range noise from a seeded generator on that device; beams with no
    return within far_plane read 0."""
    import torch

    f64 = torch.float64
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    dS = torch.tensor(lidar_directions(spec), dtype=f64, device=device).reshape(-1, 3)
    lo = torch.tensor(street.box_lo, dtype=f64, device=device)
    hi = torch.tensor(street.box_hi, dtype=f64, device=device)
    cxy = torch.tensor(street.cyl_xy, dtype=f64, device=device)
    cr = torch.tensor(street.cyl_r, dtype=f64, device=device)
    ch = torch.tensor(street.cyl_h, dtype=f64, device=device)
    inf = float("inf")
    out = torch.empty((len(poses), spec.height, spec.width), dtype=torch.float32, device=device)
    for k, T in enumerate(poses):
        T = torch.tensor(np.asarray(T, np.float64), dtype=f64, device=device)
        d = dS @ T[:3, :3].T
        o = T[:3, 3]
        # ground plane
        t = torch.where(d[:, 2] < 0, -o[2] / d[:, 2], torch.full_like(d[:, 2], inf))
        # boxes (slab test), in chunks of boxes
        inv = 1.0 / d
        for b0 in range(0, len(street.box_lo), 8):
            t1 = (lo[None, b0:b0 + 8] - o) * inv[:, None]
            t2 = (hi[None, b0:b0 + 8] - o) * inv[:, None]
            tn = torch.minimum(t1, t2).amax(-1)
            tf = torch.maximum(t1, t2).amin(-1)
            hit = (tn <= tf) & (tn > 0)
            t = torch.minimum(t, torch.where(hit, tn, torch.full_like(tn, inf)).amin(-1))
        # vertical cylinders: side wall between z = 0 and the cylinder's height
        a = d[:, 0] ** 2 + d[:, 1] ** 2
        for c0 in range(0, len(street.cyl_r), 16):
            oc = o[None, :2] - cxy[c0:c0 + 16]
            b = 2.0 * (d[:, None, 0] * oc[None, :, 0] + d[:, None, 1] * oc[None, :, 1])
            cc = (oc ** 2).sum(-1)[None] - cr[None, c0:c0 + 16] ** 2
            disc = b * b - 4.0 * a[:, None] * cc
            t0 = (-b - torch.sqrt(torch.clamp(disc, min=0.0))) / (2.0 * a[:, None])
            z = o[2] + t0 * d[:, None, 2]
            hit = (disc > 0) & (t0 > 0) & (z >= 0) & (z <= ch[None, c0:c0 + 16])
            t = torch.minimum(t, torch.where(hit, t0, torch.full_like(t0, inf)).amin(-1))
        r = t
        if noise > 0:
            r = r + torch.randn(r.shape, generator=gen, dtype=f64, device=device) * noise
        r = torch.where(torch.isfinite(t) & (t <= spec.far_plane), r, torch.zeros_like(r))
        out[k] = r.reshape(spec.height, spec.width).to(torch.float32)
    return out

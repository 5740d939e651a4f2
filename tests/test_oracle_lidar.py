"""Oracle OusterLidar model (ouster_lidar.cpp:59-253, ouster_lidar_impl.hpp:15-32).

The projection library of the reference (srl::projection::OusterLidar) is not vendored, so the
beam model is this build's own definition, stated in oracle/se8_oracle.c and shared operation for
operation with the GPU kernels.  These tests pin it by properties rather than by restating it:
the fixed-polynomial trigonometry against numpy, backProject -> project round trips over every
pixel, the frustum cone test against a float64 evaluation of the reference's formula
(asin / sin), and the TSDF of a lidar integration against the closed form of a plane.
"""
import numpy as np
import pytest

from oracle import oracle as O

W, H = 256, 32
ELEV = np.linspace(22.5, -22.5, H)


def _lidar(az=None):
    return O.OracleLidar(W, H, 0.5, 60.0, ELEV, az)


def test_atan2_and_sin_polynomials():
    rng = np.random.default_rng(0)
    y = rng.normal(size=2000).astype(np.float32)
    x = rng.normal(size=2000).astype(np.float32)
    got = np.array([O.atan2_det(a, b) for a, b in zip(y, x)])
    assert np.abs(got - np.arctan2(y.astype(np.float64), x.astype(np.float64))).max() < 3e-7
    # axes and quadrant boundaries
    for a, b, want in [(0, 1, 0), (1, 0, np.pi / 2), (0, -1, np.pi), (-1, 0, -np.pi / 2), (0, 0, 0)]:
        assert abs(O.atan2_det(a, b) - want) < 3e-7
    xs = np.linspace(0, np.pi, 3001)
    s = np.array([O.sin_det(v) for v in xs])
    assert np.abs(s - np.sin(xs)).max() < 5e-7


@pytest.mark.parametrize("az", [None, np.linspace(-1.5, 1.5, H)])
def test_backproject_project_round_trip(az):
    lid = _lidar(az)
    dirs = lid.directions()
    assert np.allclose(np.linalg.norm(dirs, axis=-1), 1.0, atol=1e-6)
    for r in (0.7, 13.0, 55.0):
        uv, ok = lid.project(dirs.reshape(-1, 3) * r)
        assert ok.all()
        px = (uv + 0.5).astype(np.int32)  # round_pixel
        u, v = np.meshgrid(np.arange(W), np.arange(H))
        np.testing.assert_array_equal(px[:, 0], u.reshape(-1))
        np.testing.assert_array_equal(px[:, 1], v.reshape(-1))
        # the sample sits on the pixel centre to a small fraction of a pixel
        assert np.abs(uv[:, 0] - u.reshape(-1)).max() < 1e-3
        assert np.abs(uv[:, 1] - v.reshape(-1)).max() < 1e-3


def test_project_rejects_outside_vertical_fov():
    lid = _lidar()
    gap = 45.0 / (H - 1)
    for e_deg, want in [(22.5 + 0.4 * gap, True), (22.5 + 0.6 * gap, False), (-22.5 - 0.4 * gap, True),
                        (-22.5 - 0.6 * gap, False), (89.0, False), (-89.0, False)]:
        e = np.radians(e_deg)
        p = np.array([np.cos(e), 0.0, np.sin(e)]) * 10
        _, ok = lid.project(p)
        assert bool(ok[0]) == want, e_deg


def _reference_sphere_in_frustum(c, r, near, far, emin, emax):
    """ouster_lidar.cpp:195-224 in float64 with the standard asin/sin."""
    n = np.linalg.norm(c)
    if n - r > far or n + r < near:
        return False
    e = np.arcsin(c[2] / n)
    for lim, outside in ((emin, e < emin), (emax, e > emax)):
        if outside and np.sin(abs(e - lim)) * n > r:
            return False
    return True


def test_sphere_in_frustum_matches_reference_formula():
    lid = _lidar()
    emin, emax = np.radians(-22.5), np.radians(22.5)
    rng = np.random.default_rng(1)
    checked = 0
    for _ in range(4000):
        c = rng.uniform(-70, 70, 3)
        c[2] = rng.uniform(-40, 40)
        r = rng.uniform(0.01, 3.0)
        want = _reference_sphere_in_frustum(c, r, 0.5, 60.0, emin, emax)
        n = np.linalg.norm(c)
        e = np.arcsin(c[2] / n)
        margins = [abs(n - r - 60.0), abs(n + r - 0.5), abs(np.sin(abs(e - emin)) * n - r),
                   abs(np.sin(abs(e - emax)) * n - r)]
        if min(margins) < 1e-4:
            continue  # float32 rounding can legitimately flip cases on the boundary
        assert lid.sphere_in_frustum(c, r) == want, (c, r)
        checked += 1
    assert checked > 3900


def test_lidar_integration_plane_closed_form():
    """A plane x = X0 seen from the origin: range along pixel ray d is X0 / d_x.  For every updated
    voxel the TSDF is min((range - |p|) / mu, 1) up to the pixel quantisation of the ray (the
    voxel's own ray and its pixel's ray differ by at most half a pixel in each angle)."""
    lid = _lidar()
    res, mu = 0.05, 0.2
    x0 = 6.0
    dirs = lid.directions(precision="f64")
    with np.errstate(divide="ignore"):
        rng_img = np.where(dirs[..., 0] > 0.05, x0 / dirs[..., 0], 0.0).astype(np.float32)
    m = O.OracleMap(25.6, res, mu_factor=mu / res)
    stats = m.integrate_lidar(0, rng_img, lid, np.eye(4))
    assert stats["new"] > 0
    keys, ts, tsdf, w = m.dump_blocks()
    assert (ts == 0).all()
    upd = w > 0
    coords = np.array([O.decode_key(int(k))[0] for k in keys])  # block corner voxels
    off = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).transpose(2, 1, 0, 3)
    vox = coords[:, None, :] + off.reshape(-1, 3)[None]  # x-fastest voxel order within a block
    p = (vox + 0.5) * res - 12.8  # voxel centres in the world (= sensor) frame, t_MW = dim / 2
    n = np.linalg.norm(p, axis=-1)
    want = np.minimum((x0 * n / p[..., 0] - n) / mu, 1.0)
    # half a pixel of angle at range R moves the plane hit by about R * dangle * tan(incidence)
    dang = max(2 * np.pi / W, np.radians(45.0 / (H - 1))) / 2
    tol = (n * dang * (1 + np.abs(p[..., 1:]).max(-1) / np.maximum(p[..., 0], 1e-3)) + 1e-4) / mu
    err = np.abs(tsdf - want)
    assert upd.sum() > 10000
    assert (err[upd] <= tol[upd]).mean() > 0.99
    # the zero crossing lies at the plane: where the signed distance exceeds twice the quantisation
    # bound, the sign of the TSDF is the side of the plane
    sdf_true = x0 * n / p[..., 0] - n
    far_from_plane = upd & (np.abs(sdf_true) > 2 * tol * mu) & (np.abs(tsdf) < 0.99)
    assert far_from_plane.sum() > 1000
    np.testing.assert_array_equal(np.sign(tsdf[far_from_plane]), np.sign(x0 - p[..., 0][far_from_plane]))

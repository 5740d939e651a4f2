"""The oracle's tracker pieces against closed forms: 6x6 solve vs numpy, the SE(3) exponential
vs scipy's matrix exponential, the reduction's b = sum(e J) vs -grad(0.5 sum e^2) by FP64 central
differences, pyramid KATs, and end-to-end convergence to ground truth."""
import numpy as np
import pytest
import scipy.linalg

from oracle import oracle as O
from tests.helpers import CAM


def hat(w):
    return np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])


def test_solve_matches_numpy():
    rng = np.random.default_rng(0)
    for _ in range(50):
        A = rng.normal(size=(6, 6))
        M = A @ A.T + 0.5 * np.eye(6)
        b = rng.normal(size=6)
        vals = np.concatenate([b, M[np.triu_indices(6)]])  # row-major upper triangle
        x = O.icp_solve(vals, "f64")
        np.testing.assert_allclose(x, np.linalg.solve(M, b), rtol=1e-9, atol=1e-12)
        x32 = O.icp_solve(vals.astype(np.float32), "f32")
        np.testing.assert_allclose(x32, np.linalg.solve(M, b), rtol=2e-3, atol=2e-3)
    # a non positive-definite system gives x = 0 (icp.cpp:37-40)
    M = np.diag([1.0, 1.0, 1.0, 1.0, 1.0, -1.0])
    vals = np.concatenate([np.ones(6), M[np.triu_indices(6)]])
    np.testing.assert_array_equal(O.icp_solve(vals, "f64"), np.zeros(6))


def test_exp_matches_matrix_exponential():
    rng = np.random.default_rng(1)
    for scale in (1e-12, 1e-3, 0.1, 1.0):
        for _ in range(10):
            a = rng.normal(size=6) * scale
            T = O.se3_exp(a, "f64")
            X = np.zeros((4, 4))
            X[:3, :3] = hat(a[3:])
            X[:3, 3] = a[:3]
            np.testing.assert_allclose(T, scipy.linalg.expm(X), atol=1e-12)
            T32 = O.se3_exp(a.astype(np.float32), "f32")
            np.testing.assert_allclose(T32, scipy.linalg.expm(X), atol=5e-6)


def _plane_scene(T_ref, tilt=0.15, D=2.0):
    """Reference maps (points/normals) of the plane z = D + tilt*x (world) raycast from T_ref, and
    the input vertex/normal maps of the same plane seen from T_ref (sensor frame)."""
    H, W = CAM.height, CAM.width
    v, u = np.mgrid[0:H, 0:W].astype(np.float64)
    dC = np.stack([(u - CAM.cx) / CAM.fx, (v - CAM.cy) / CAM.fy, np.ones_like(u)], -1)
    n = np.array([-tilt, 0.0, 1.0])
    n /= np.linalg.norm(n)
    R, o = T_ref[:3, :3], T_ref[:3, 3]
    dW = dC @ R.T
    t = (D - o[2] + tilt * o[0]) / (dW[..., 2] - tilt * dW[..., 0])
    ref_pts = o + dW * t[..., None]
    ref_nrm = np.broadcast_to(-n, ref_pts.shape).copy()  # towards the camera
    in_pts = dC * t[..., None]
    in_nrm = np.broadcast_to(-(R.T @ n), in_pts.shape).copy()
    return ref_pts, ref_nrm, in_pts, in_nrm


def test_reduction_b_is_minus_gradient_of_half_squared_error():
    T_ref = np.eye(4)
    ref_pts, ref_nrm, in_pts, in_nrm = _plane_scene(T_ref)
    pin = O.pinhole()
    rng = np.random.default_rng(2)
    xi0 = np.concatenate([rng.normal(0, 0.004, 3), rng.normal(0, 0.002, 3)])
    T0 = O.se3_exp(xi0, "f64") @ T_ref
    _, _, _, sums = O.icp_track_reduce(in_pts, in_nrm, ref_pts, ref_nrm, pin, T0, T_ref, precision="f64")
    b = sums[1:7]
    n0 = sums[28]

    def cost(xi):
        T = O.se3_exp(xi, "f64") @ T0
        _, _, _, s = O.icp_track_reduce(in_pts, in_nrm, ref_pts, ref_nrm, pin, T, T_ref, precision="f64")
        assert s[28] == n0  # same association
        return 0.5 * s[0]

    eps = 1e-7
    grad = np.array([(cost(eps * e) - cost(-eps * e)) / (2 * eps) for e in np.eye(6)])
    np.testing.assert_allclose(b, -grad, rtol=1e-4, atol=1e-9 * np.abs(b).max())


def test_half_sample_kat():
    d = np.array([[1.0, 1.1, 5.0, 5.0],
                  [1.2, 9.0, 5.1, 0.0],
                  [2.0, 2.0, 3.0, 3.2],
                  [2.0, 2.4, 3.1, 3.0]], np.float32)
    out = O.half_sample(d)
    # pixels within 0.3 of the block's top-left pixel are averaged in the order (0,0),(1,0),(0,1),(1,1)
    want = np.array([[np.float32(np.float32(np.float32(1.0) + np.float32(1.1)) + np.float32(1.2)) / 3,
                      np.float32(np.float32(5.0 + 5.0) + np.float32(5.1)) / 3],
                     [np.float32(2.0 + 2.0 + 2.0) / 3,
                      np.float32(np.float32(np.float32(3.0) + np.float32(3.2)) + np.float32(3.1)) + np.float32(3.0)]],
                    np.float32)
    want[1, 1] = want[1, 1] / np.float32(4)
    np.testing.assert_array_equal(out, want)


def test_depth_to_points_and_normals_of_a_plane():
    pin = O.pinhole()
    H, W = 480, 640
    v, u = np.mgrid[0:H, 0:W].astype(np.float64)
    depth = np.full((H, W), 2.0, np.float32)
    depth[10, 10] = 0.0
    for level in (0, 1, 2):
        dl = depth[:: 1 << level, :: 1 << level].copy()
        pts = O.depth_to_points(dl, pin, level, "f64")
        s = 1 << level
        fx = 525.0 / s
        cx = (319.5 + 0.5) / s - 0.5
        vv, uu = np.mgrid[0:H // s, 0:W // s].astype(np.float64)
        want_x = 2.0 * (uu - cx) / fx
        ok = dl > 0
        np.testing.assert_allclose(pts[..., 0][ok], want_x[ok], atol=1e-12)
        assert np.all(pts[~ok] == 0)
        nrm = O.points_to_normals(pts, False, "f64")
        good = np.any(nrm != 0, -1)
        # a fronto-parallel plane: (right - left) x (row below - row above) points away from the
        # camera (+z); the tracker negates the raycast normals to match (icp_impl.hpp:63-69)
        np.testing.assert_allclose(nrm[good], np.broadcast_to([0, 0, 1.0], nrm[good].shape), atol=1e-12)
        # the zero pixel invalidates itself and its 4 neighbours
        if level == 0:
            assert not good[10, 10] and not good[10, 9] and not good[9, 10] and not good[10, 11] and not good[11, 10]


@pytest.fixture(scope="module")
def mapped_scene():
    from tests.helpers import spline_frames

    depth, poses = spline_frames(6, noise=False)
    m = O.OracleMap(10.24, 0.01)
    pin = O.pinhole()
    for k in range(5):
        m.integrate(k, depth[k], pin, poses[k])
    pts, nrm, _ = m.raycast(pin, poses[4])
    return depth, poses, pts, nrm, pin


def test_track_identity_keeps_pose(mapped_scene):
    depth, poses, pts, nrm, pin = mapped_scene
    ok, T, sums, iters = O.track(depth[4], pin, poses[4], pts, nrm)
    assert ok
    # the fused model and the frame differ by TSDF discretisation; ICP stays within 1 mm
    np.testing.assert_allclose(T, poses[4], atol=1e-3)


def test_track_converges_to_ground_truth(mapped_scene):
    depth, poses, pts, nrm, pin = mapped_scene
    ok, T, sums, iters = O.track(depth[5], pin, poses[4], pts, nrm)
    assert ok
    assert np.abs(T[:3, 3] - poses[5][:3, 3]).max() < 2e-3
    R_err = T[:3, :3].T @ poses[5][:3, :3]
    assert np.degrees(np.arccos(np.clip((np.trace(R_err) - 1) / 2, -1, 1))) < 0.1


def test_track_failure_restores_pose(mapped_scene):
    """checkPoseKernel (icp.cpp:227-246): too few inliers restores the previous pose."""
    depth, poses, pts, nrm, pin = mapped_scene
    far = poses[4].copy()
    far[:3, 3] += np.array([1.0, 0.0, 0.0])  # a 1 m jump fails the distance gate everywhere
    ok, T, sums, _ = O.track(depth[5], pin, far, pts, nrm)
    assert not ok
    np.testing.assert_array_equal(T, far.astype(np.float32))
    ok, T, _, _ = O.track(np.zeros_like(depth[5]), pin, poses[4], pts, nrm)
    assert not ok

"""Differentiation workload on the GPU: 64 candidate poses around a query frame's ground truth
(t ~ N(0, 3 cm), r ~ N(0, 2 deg), seeded).  The adjoint of the point-to-plane cost is checked
against the reference tracker's J^T r, whose per-pixel Jacobian is checked against FP64 central
differences; the adjoint of the raycast depth against the FP64 oracle's implicit-function gradient
and FP64 central differences (eps 1e-5); then 20 batched Gauss-Newton steps bring the candidates
to the minimum of the cost, which is the reference tracker's fixed point."""
import numpy as np
import pytest

from oracle import oracle as O
from supereight2_amd import synth
from tests.helpers import CAM, circle_frames

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

N_POSES = 64


def _sensor(se):
    return se.PinholeCamera(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)


def _pin():
    return O.pinhole(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)


@pytest.fixture(scope="module")
def scene():
    from supereight2_amd import diff, se

    depth, poses = circle_frames(50, 6)
    sensor = _sensor(se)
    gmap = se.TSDFMap(10.24, 0.01)
    integ = se.MapIntegrator(gmap)
    o64 = O.OracleMap(10.24, 0.01, precision="f64")
    pin = _pin()
    for k in range(5):
        print(f"integrate {k}", flush=True)
        integ.integrateDepth(k, torch.from_numpy(depth[k]).cuda(), sensor, poses[k])
        o64.integrate(k, depth[k], pin, poses[k].astype(np.float64))
    q = 5
    T_gt = poses[q].astype(np.float64)
    ref_p = torch.zeros((CAM.height, CAM.width, 3), dtype=torch.float32, device="cuda")
    ref_n = torch.zeros_like(ref_p)
    se.raycast_volume(gmap, sensor, T_gt, ref_p, ref_n)
    rng = np.random.default_rng(11)
    twists = np.concatenate([rng.normal(0, 0.03, (N_POSES, 3)), rng.normal(0, np.radians(2.0), (N_POSES, 3))], 1)
    cands = np.stack([diff.se3_exp(tw) @ T_gt for tw in twists])
    tracker = se.Tracker(gmap, sensor)
    d = torch.from_numpy(depth[q]).cuda()
    return dict(se=se, diff=diff, gmap=gmap, o64=o64, pin=pin, sensor=sensor, depth=depth[q], d=d, T_gt=T_gt,
                ref_p=ref_p, ref_n=ref_n, cands=cands, tracker=tracker)


def test_cost_gradient_is_reference_jtr_and_fd(scene):
    s = scene
    cost, grad, jtj, inl = s["diff"].icp_cost_grad(s["tracker"], s["d"], s["ref_p"], s["ref_n"], s["T_gt"], s["cands"])
    assert np.all(inl > 1000)
    # the reference reduction (trackKernel + reduceKernel restated in the oracle) on the same inputs
    pin = s["pin"]
    in_pts = O.depth_to_points(s["depth"], pin, 0)
    in_nrm = O.points_to_normals(in_pts, False)
    rp, rn = s["ref_p"].cpu().numpy(), s["ref_n"].cpu().numpy()
    in64 = O.depth_to_points(s["depth"], pin, 0, "f64")
    nr64 = O.points_to_normals(in64, False, "f64")
    for k in range(0, N_POSES, 8):
        _, _, _, sums = O.icp_track_reduce(in_pts, in_nrm, rp, rn, pin, s["cands"][k].astype(np.float32),
                                           s["T_gt"].astype(np.float32))
        assert sums[28] == inl[k]
        np.testing.assert_allclose(grad[k], -sums[1:7], rtol=2e-3, atol=2e-3 * np.abs(sums[1:7]).max())
        np.testing.assert_allclose(cost[k], 0.5 * sums[0], rtol=2e-3)
    # FP64 central differences, per pixel.  The cost is piecewise smooth: a pixel whose projective
    # association or inlier test flips between T - eps and T + eps contributes a jump, and at
    # 640x480 about eps * 500 px/rad of the pixels sit that close to a pixel border.  Such pixels
    # are found without using J, by their second difference (a smooth residual has
    # |e+ + e- - 2 e0| = O(eps^2)), and left out of the comparison.
    rp64, rn64 = rp.astype(np.float64), rn.astype(np.float64)
    for k in (0, 17, 42):
        T0 = s["cands"][k].astype(np.float64)

        def track64(xi):
            res, err, J, _ = O.icp_track_reduce(in64, nr64, rp64, rn64, pin, O.se3_exp(xi, "f64") @ T0, s["T_gt"],
                                                precision="f64")
            return res == 1, err, J

        ok0, e0, J0 = track64(np.zeros(6))
        eps = 1e-6
        de = np.zeros(e0.shape + (6,))
        smooth = ok0.copy()
        for a in range(6):
            xi = np.zeros(6)
            xi[a] = eps
            (okp, ep, _), (okm, em, _) = track64(xi), track64(-xi)
            de[..., a] = (ep - em) / (2 * eps)
            smooth &= okp & okm & (np.abs(ep + em - 2 * e0) < 1e-9)
        assert smooth.sum() > 0.98 * ok0.sum()
        # de/dxi = -J on every smooth pixel (left perturbation T <- exp(xi) T, xi = [v; w])
        rel = np.linalg.norm(de[smooth] + J0[smooth], axis=-1) / np.linalg.norm(J0[smooth], axis=-1)
        assert np.median(rel) < 1e-6 and np.max(rel) < 1e-3
        fd = (e0[smooth][:, None] * de[smooth]).sum(0)
        adj = -(e0[smooth][:, None] * J0[smooth]).sum(0)
        print("pose", k, "smooth pixels", smooth.sum(), "of", ok0.sum(), "fd64", fd, "adjoint64", adj)
        np.testing.assert_allclose(adj, fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())


def test_autograd_function_matches_kernel(scene):
    s = scene
    delta = torch.zeros((N_POSES, 6), dtype=torch.float64, requires_grad=True)
    cost = s["diff"].point_to_plane_cost(delta, s["tracker"], s["d"], s["ref_p"], s["ref_n"], s["T_gt"], s["cands"])
    cost.sum().backward()
    _, grad, _, _ = s["diff"].icp_cost_grad(s["tracker"], s["d"], s["ref_p"], s["ref_n"], s["T_gt"], s["cands"])
    np.testing.assert_allclose(delta.grad.numpy(), grad, rtol=1e-6, atol=1e-9)


def test_depth_gradient_vs_fp64_oracle(scene):
    s = scene
    poses = s["cands"][:4]
    depth, grad, valid = s["diff"].raycast_depth_grad(s["gmap"], s["sensor"], poses)
    depth, grad, valid = depth.cpu().numpy(), grad.cpu().numpy(), valid.cpu().numpy()
    assert valid.mean() > 0.5
    for k in range(2):
        d64, g64, v64 = s["o64"].raycast_depth_grad(s["pin"], poses[k])
        both = valid[k] & v64
        assert both.mean() > 0.5
        # fp32 GPU map vs FP64 oracle map: depths agree to a fraction of a voxel
        assert np.median(np.abs(depth[k][both] - d64[both])) < 5e-4
        rel = np.linalg.norm(grad[k][both] - g64[both], axis=-1) / np.linalg.norm(g64[both], axis=-1)
        print("pose", k, "median rel grad diff", np.median(rel))
        assert np.median(rel) < 2e-2
        # FP64 central differences of the oracle depth confirm the FP64 gradient on the same pixels
        eps = 1e-5
        fd = np.zeros_like(g64)
        ok = v64.copy()
        for a in range(6):
            xi = np.zeros(6)
            xi[a] = eps
            dp, _, vp = s["o64"].raycast_depth_grad(s["pin"], O.se3_exp(xi, "f64") @ poses[k])
            dm, _, vm = s["o64"].raycast_depth_grad(s["pin"], O.se3_exp(-xi, "f64") @ poses[k])
            fd[..., a] = (dp - dm) / (2 * eps)
            ok &= vp & vm
        sel = ok & valid[k]
        rel_fd = np.linalg.norm(grad[k][sel] - fd[sel], axis=-1) / np.linalg.norm(fd[sel], axis=-1)
        print("pose", k, "median rel diff to FP64 FD", np.median(rel_fd))
        assert np.median(rel_fd) < 2e-2


def _gauss_newton(s, T, steps):
    T = T.copy()
    for _ in range(steps):
        _, grad, jtj, _ = s["diff"].icp_cost_grad(s["tracker"], s["d"], s["ref_p"], s["ref_n"], s["T_gt"], T)
        for k in range(len(T)):
            A = np.zeros((6, 6))
            A[np.triu_indices(6)] = jtj[k]
            A = A + A.T - np.diag(np.diag(A))
            step = np.linalg.solve(A.astype(np.float64), -grad[k].astype(np.float64))
            T[k] = s["diff"].se3_exp(step) @ T[k]
    return T


def _rot_deg(A, B):
    # the angle of A^T B from its antisymmetric part (synth.rotation_angle_deg): the arccos of the
    # trace cannot resolve 0.01 degrees between a float32 pose and a float64 one
    return synth.rotation_angle_deg(np.asarray(A, np.float64)[:3, :3].T @ np.asarray(B, np.float64)[:3, :3])


def test_batched_gauss_newton_refinement_converges(scene):
    """The candidates descend to the minimum of the point-to-plane cost.  That minimum is the pose
    the reference tracker converges to from the ground truth; it sits a few mm from the ground truth
    because the model is a 5-frame fusion of noisy depth."""
    s = scene
    T_star = _gauss_newton(s, s["T_gt"][None], 20)[0]
    ok, T_trk, _ = s["tracker"].track(s["d"], s["T_gt"], s["ref_p"], s["ref_n"])
    assert ok
    off_trk = np.linalg.norm(T_trk[:3, 3] - T_star[:3, 3])
    off_gt = np.linalg.norm(T_star[:3, 3] - s["T_gt"][:3, 3])
    print(f"cost minimum: {off_gt * 1000:.3f} mm from ground truth, {off_trk * 1000:.4f} mm from the tracker's pose")
    assert off_trk < 1e-4 and _rot_deg(T_trk, T_star) < 0.01
    assert off_gt < 5e-3
    T = _gauss_newton(s, s["cands"], 20)
    err0 = np.linalg.norm(s["cands"][:, :3, 3] - T_star[:3, 3], axis=1)
    err = np.linalg.norm(T[:, :3, 3] - T_star[:3, 3], axis=1)
    rot = np.array([_rot_deg(T[k], T_star) for k in range(N_POSES)])
    print(f"initial distance median {np.median(err0) * 100:.2f} cm; after 20 steps median {np.median(err) * 1000:.4f} mm,"
          f" {np.mean(err < 1e-3) * 100:.0f}% within 1 mm, rot median {np.median(rot):.5f} deg")
    assert np.median(err) < 1e-4 and np.median(rot) < 1e-3
    assert np.mean(err < 1e-3) > 0.7

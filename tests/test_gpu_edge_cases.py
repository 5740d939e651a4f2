"""Edge cases of the integration path on the GPU against the oracle.

- invalid depth values (NaN, +-inf, negative, 0, below the near plane, beyond the far plane) mixed
  into two tracking frames: allocation, timestamps, weights and octree bit-exact, TSDF values equal
  (NaN where the reference's std::min(sdf / mu, 1) keeps a NaN), raycast equal;
- an all-invalid frame (no block allocated, every ray a miss), then a valid frame;
- a map whose block pool is too small: the overflow is reported as SE8_ERR_CAPACITY at the next
  integration, the map keeps the blocks that fit, stays readable, and refuses further integration;
- tracking against empty reference maps (the check fails and the pose is restored bit for bit);
- an odd 322 x 243 camera through integration, raycast and tracking.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import CAM, spline_frames

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _sensor(se):
    return se.PinholeCamera(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)


def _pin():
    return O.pinhole(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)


def _corrupt(depth, seed):
    d = depth.copy()
    r = np.random.default_rng(seed).random(d.shape)
    for i, v in enumerate([np.nan, np.inf, -np.inf, -1.0, 0.0, 0.05, 9.0]):
        d[(r >= 0.01 * i) & (r < 0.01 * (i + 1))] = v
    return d.astype(np.float32)


def _compare_maps(gmap, omap):
    gk, gts, gt, gw = gmap.dump_blocks()
    ok, ots, ot, ow = omap.dump_blocks()
    np.testing.assert_array_equal(gk, ok)
    np.testing.assert_array_equal(gts, ots)
    np.testing.assert_array_equal(gw, ow)
    np.testing.assert_array_equal(np.isnan(gt), np.isnan(ot))
    fin = ~np.isnan(gt)
    np.testing.assert_array_equal(gt[fin].view(np.uint32), ot[fin].view(np.uint32))
    for a, b in zip(gmap.dump_nodes(), omap.dump_nodes()):
        np.testing.assert_array_equal(a, b)
    return int(np.isnan(gt).sum())


def test_invalid_depth_values_match_oracle():
    from supereight2_amd import se

    depth, poses = spline_frames(2)
    sensor, pin = _sensor(se), _pin()
    gmap, omap = se.TSDFMap(10.24, 0.01), O.OracleMap(10.24, 0.01)
    integ = se.MapIntegrator(gmap)
    for k in range(2):
        d = _corrupt(depth[k], 100 + k)
        print(f"frame {k}: {np.isnan(d).sum()} NaN, {np.isinf(d).sum()} inf pixels", flush=True)
        integ.integrateDepth(k, torch.from_numpy(d).cuda(), sensor, poses[k])
        omap.integrate(k, d, pin, poses[k])
    torch.cuda.synchronize()
    n_nan = _compare_maps(gmap, omap)
    print("NaN voxels", n_nan)
    assert n_nan > 0  # NaN depth samples reach the TSDF as in the reference
    pts = torch.zeros((CAM.height, CAM.width, 3), dtype=torch.float32, device="cuda")
    nrm = torch.zeros_like(pts)
    se.raycast_volume(gmap, sensor, poses[1], pts, nrm)
    op, on, _ = omap.raycast(pin, poses[1])
    gp, gn = pts.cpu().numpy(), nrm.cpu().numpy()
    np.testing.assert_array_equal(gp.reshape(op.shape), op)
    np.testing.assert_array_equal(gn.reshape(on.shape), on)
    gmap.close()


def test_all_invalid_frame_then_valid_frame():
    from supereight2_amd import se

    depth, poses = spline_frames(2)
    sensor, pin = _sensor(se), _pin()
    gmap, omap = se.TSDFMap(10.24, 0.01), O.OracleMap(10.24, 0.01)
    integ = se.MapIntegrator(gmap)
    empty = np.zeros_like(depth[0])
    integ.integrateDepth(0, torch.from_numpy(empty).cuda(), sensor, poses[0])
    omap.integrate(0, empty, pin, poses[0])
    torch.cuda.synchronize()
    assert gmap.info().num_blocks == 0 and omap.num_blocks == 0
    pts = torch.ones((CAM.height, CAM.width, 3), dtype=torch.float32, device="cuda")
    nrm = torch.ones_like(pts)
    se.raycast_volume(gmap, sensor, poses[0], pts, nrm)
    assert float(pts.abs().sum()) == 0.0 and float(nrm.abs().sum()) == 0.0
    integ.integrateDepth(1, torch.from_numpy(depth[1]).cuda(), sensor, poses[1])
    omap.integrate(1, depth[1], pin, poses[1])
    torch.cuda.synchronize()
    assert _compare_maps(gmap, omap) == 0
    gmap.close()


def test_capacity_overflow_is_reported_and_map_stays_readable():
    from supereight2_amd import capi, se

    depth, poses = spline_frames(2)
    sensor = _sensor(se)
    cap = 512
    gmap = se.TSDFMap(10.24, 0.01, max_blocks=cap)
    integ = se.MapIntegrator(gmap)
    integ.integrateDepth(0, torch.from_numpy(depth[0]).cuda(), sensor, poses[0])  # needs ~25k blocks
    torch.cuda.synchronize()
    for k in (1, 2):
        with pytest.raises(capi.Se8Error, match="capacity"):
            integ.integrateDepth(k, torch.from_numpy(depth[1]).cuda(), sensor, poses[1])
    torch.cuda.synchronize()
    info = gmap.info()
    assert info.num_blocks == cap
    keys, ts, tsdf, w = gmap.dump_blocks()
    assert len(keys) == cap and (ts == 0).all()
    pts = torch.zeros((CAM.height, CAM.width, 3), dtype=torch.float32, device="cuda")
    nrm = torch.zeros_like(pts)
    se.raycast_volume(gmap, sensor, poses[0], pts, nrm)
    torch.cuda.synchronize()
    assert np.isfinite(pts.cpu().numpy()).all()
    gmap.close()


def test_track_failure_restores_pose():
    """Against empty reference maps no pixel associates: the check fails (icp.cpp:227-246) and the
    pose is restored bit for bit, as in the oracle."""
    from supereight2_amd import se

    depth, poses = spline_frames(2)
    sensor, pin = _sensor(se), _pin()
    gmap = se.TSDFMap(10.24, 0.01)
    tracker = se.Tracker(gmap, sensor)
    zeros = torch.zeros((CAM.height, CAM.width, 3), dtype=torch.float32, device="cuda")
    T0 = poses[0].copy()
    ok_g, T_g, sums_g = tracker.track(torch.from_numpy(depth[1]).cuda(), T0, zeros, zeros)
    z = np.zeros((CAM.height, CAM.width, 3), np.float32)
    ok_o, T_o, sums_o, _ = O.track(depth[1], pin, T0, z, z)
    assert not ok_g and not ok_o
    np.testing.assert_array_equal(T_g.view(np.uint32), T0.view(np.uint32))
    assert sums_g[28] == sums_o[28] == 0
    tracker.close()
    gmap.close()


def test_odd_image_size_pipeline_matches_oracle():
    """A 322 x 243 camera (odd sizes at every pyramid level, partial raycast tiles): integration and
    raycast bit-exact, tracking within the tolerance of the 640 x 480 test."""
    from supereight2_amd import se, synth

    cam = synth.Camera(width=322, height=243, fx=264.0, fy=264.0, cx=160.5, cy=121.0, near_plane=0.4, far_plane=6.0)
    scene = synth.make_scene(0)
    poses = np.stack(synth.spline_trajectory(300, seed=1)[:5]).astype(np.float32)
    rng = np.random.default_rng(7)
    depth = np.stack([synth.render_depth(scene, cam, T, rng) for T in poses])
    sensor = se.PinholeCamera(cam.width, cam.height, cam.fx, cam.fy, cam.cx, cam.cy, cam.near_plane, cam.far_plane)
    pin = O.pinhole(cam.width, cam.height, cam.fx, cam.fy, cam.cx, cam.cy, cam.near_plane, cam.far_plane)
    gmap, omap = se.TSDFMap(10.24, 0.01), O.OracleMap(10.24, 0.01)
    integ = se.MapIntegrator(gmap)
    for k in range(4):
        integ.integrateDepth(k, torch.from_numpy(depth[k]).cuda(), sensor, poses[k])
        omap.integrate(k, depth[k], pin, poses[k])
    torch.cuda.synchronize()
    assert _compare_maps(gmap, omap) == 0
    pts = torch.zeros((cam.height, cam.width, 3), dtype=torch.float32, device="cuda")
    nrm = torch.zeros_like(pts)
    se.raycast_volume(gmap, sensor, poses[3], pts, nrm)
    op, on, _ = omap.raycast(pin, poses[3])
    np.testing.assert_array_equal(pts.cpu().numpy(), op)
    np.testing.assert_array_equal(nrm.cpu().numpy(), on)
    tracker = se.Tracker(gmap, sensor)
    ok_g, T_g, _ = tracker.track(torch.from_numpy(depth[4]).cuda(), poses[3], pts, nrm)
    ok_o, T_o, _, _ = O.track(depth[4], pin, poses[3], op, on)
    print("odd size track ok", ok_g, ok_o)
    assert ok_g == ok_o
    np.testing.assert_allclose(T_g, T_o, atol=1e-4)
    tracker.close()
    gmap.close()


def test_optional_outputs_are_validated():
    """The optional output tensors (raycast scale, the downsample index map, track_device's ok) are
    checked for device, dtype, contiguity and size before any kernel is launched."""
    from supereight2_amd import se

    sensor = _sensor(se)
    H, W = CAM.height, CAM.width
    gmap = se.TSDFMap(10.24, 0.01)
    pts = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    nrm = torch.zeros_like(pts)
    T = np.eye(4, dtype=np.float32)
    bad_scales = [torch.zeros((H, W), dtype=torch.int32, device="cuda"),   # wrong dtype
                  torch.zeros((H, W - 1), dtype=torch.int8, device="cuda"),  # short
                  torch.zeros((H, W), dtype=torch.int8),                     # host memory
                  torch.zeros((W, 2 * H), dtype=torch.int8, device="cuda")[:, ::2]]  # not contiguous
    for sc in bad_scales:
        with pytest.raises(ValueError, match="scale"):
            se.raycast_volume(gmap, sensor, T, pts, nrm, scale=sc)
    se.raycast_volume(gmap, sensor, T, pts, nrm, scale=torch.zeros((H, W), dtype=torch.int8, device="cuda"))
    depth = torch.ones((64, 64), dtype=torch.float32, device="cuda")
    out = torch.zeros((32, 32), dtype=torch.float32, device="cuda")
    for mp in (torch.zeros((32, 32), dtype=torch.int32, device="cuda"),
               torch.zeros((32, 31), dtype=torch.int64, device="cuda")):
        with pytest.raises(ValueError, match="downsample_map"):
            se.downsample_depth(depth, out, mp)
    tr = se.Tracker(gmap, sensor)
    T_dev = torch.zeros((2, 16), dtype=torch.float32, device="cuda")
    d = torch.zeros((H, W), dtype=torch.float32, device="cuda")
    for ok in (torch.zeros(1, dtype=torch.int64, device="cuda"), torch.zeros(1, dtype=torch.int32),
               torch.zeros(2, dtype=torch.int32, device="cuda")):
        with pytest.raises(ValueError, match="ok"):
            tr.track_device(d, T_dev[0], T_dev[1], pts, nrm, ok=ok)
    torch.cuda.synchronize()
    tr.close()
    gmap.close()

"""Tracker::track on the GPU against the oracle beyond a single frame, through the C-ABI.

* A 30-frame tracked sequence (the spline workload's first frames): at every frame the oracle tracks
  from the same previous pose against the same raycast, then both integrate and raycast at the
  GPU's tracked pose, so the maps stay bit-identical and every frame's tracking starts from equal
  inputs.  Bar: check outcome equal, pose within 1e-4, raycast and final voxels bit-exact.  Run for
  the FP32 path (plain reduction) and the FP16-TSDF + compensated-ICP path (the oracle with FP16
  storage and the reference reduction).
* One Gauss-Newton pass at a common pose, per pyramid level: the inlier and failure counters are
  equal, the sums agree to summation order.
* The tracker's input pyramid (robust half-sampling, depth_to_point_cloud, point_cloud_to_normal)
  bit-exact at every level, for right- and left-handed sensors, odd image sizes and 1 to 6 levels.
* Trackers with other level counts against the oracle with the same configuration.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import CAM, spline_frames

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SEQ_FRAMES = 30
SEQ_SEED = 10  # the benchmark's sequence: the reference tracker holds along it


def _sensor(se, cam=CAM, w=None, h=None, fy_sign=1.0):
    return se.PinholeCamera(w or cam.width, h or cam.height, cam.fx, fy_sign * cam.fy, cam.cx, cam.cy,
                            cam.near_plane, cam.far_plane)


def _pin(w=None, h=None, fy_sign=1.0):
    return O.pinhole(w or CAM.width, h or CAM.height, CAM.fx, fy_sign * CAM.fy, CAM.cx, CAM.cy, CAM.near_plane,
                     CAM.far_plane)


def _raycast_both(se, gmap, omap, sensor, pin, T):
    pts = torch.zeros((CAM.height, CAM.width, 3), dtype=torch.float32, device="cuda")
    nrm = torch.zeros_like(pts)
    se.raycast_volume(gmap, sensor, T, pts, nrm)
    op, on, _ = omap.raycast(pin, T)
    np.testing.assert_array_equal(pts.cpu().numpy(), op)
    np.testing.assert_array_equal(nrm.cpu().numpy(), on)
    return pts, nrm, op, on


@pytest.mark.parametrize("storage,compensated", [("f32", False), ("f16", True)])
def test_tracked_sequence_matches_oracle(storage, compensated):
    from supereight2_amd import se

    depth, gt = spline_frames(SEQ_FRAMES + 1, seed=SEQ_SEED)
    sensor, pin = _sensor(se), _pin()
    gmap = se.TSDFMap(10.24, 0.01, storage=storage)
    integ = se.MapIntegrator(gmap)
    omap = O.OracleMap(10.24, 0.01, storage=storage)
    tracker = se.Tracker(gmap, sensor, se.TrackerConfig(compensated=compensated))
    T = gt[0].astype(np.float32)
    integ.integrateDepth(0, torch.from_numpy(depth[0]).cuda(), sensor, T)
    omap.integrate(0, depth[0], pin, T)
    pts, nrm, op, on = _raycast_both(se, gmap, omap, sensor, pin, T)
    worst_t = worst_r = 0.0
    n_ok = 0
    traj = [T]
    for k in range(1, SEQ_FRAMES + 1):
        ok_g, T_g, sums_g = tracker.track(torch.from_numpy(depth[k]).cuda(), T, pts, nrm)
        ok_o, T_o, sums_o, iters = O.track(depth[k], pin, T, op, on)
        dt = float(np.abs(T_g[:3, 3] - T_o[:3, 3]).max())
        dr = float(np.abs(T_g[:3, :3] - T_o[:3, :3]).max())
        worst_t, worst_r = max(worst_t, dt), max(worst_r, dr)
        print(f"frame {k}: ok {ok_g}/{ok_o} iterations {iters} |dt| {dt:.2e} |dR| {dr:.2e} "
              f"inliers {sums_g[28]:.0f}/{sums_o[28]:.0f}", flush=True)
        assert ok_g == ok_o
        assert dt < 1e-4 and dr < 1e-4
        n_ok += ok_g
        T = T_g.astype(np.float32)
        traj.append(T)
        integ.integrateDepth(k, torch.from_numpy(depth[k]).cuda(), sensor, T)
        omap.integrate(k, depth[k], pin, T)
        pts, nrm, op, on = _raycast_both(se, gmap, omap, sensor, pin, T)
    gk, _, gtsdf, gw = gmap.dump_blocks()
    ok_, _, otsdf, ow = omap.dump_blocks()
    np.testing.assert_array_equal(gk, ok_)
    np.testing.assert_array_equal(gw, ow)
    np.testing.assert_array_equal(gtsdf, otsdf)
    ate = float(np.sqrt(np.mean([np.sum((a[:3, 3] - b[:3, 3]) ** 2) for a, b in zip(traj, gt)])))
    print(f"{storage} compensated={compensated}: {n_ok}/{SEQ_FRAMES} frames pass the check, worst pose difference "
          f"{worst_t:.2e} m / {worst_r:.2e}, ATE {ate * 1000:.2f} mm, {len(gk)} blocks")
    assert n_ok == SEQ_FRAMES
    assert ate < 0.01
    tracker.close()
    gmap.close()


@pytest.fixture(scope="module")
def frame_and_model():
    from supereight2_amd import se

    depth, gt = spline_frames(6)
    sensor, pin = _sensor(se), _pin()
    gmap = se.TSDFMap(10.24, 0.01)
    integ = se.MapIntegrator(gmap)
    omap = O.OracleMap(10.24, 0.01)
    for k in range(5):
        integ.integrateDepth(k, torch.from_numpy(depth[k]).cuda(), sensor, gt[k])
        omap.integrate(k, depth[k], pin, gt[k])
    pts, nrm, op, on = _raycast_both(se, gmap, omap, sensor, pin, gt[4])
    yield dict(se=se, depth=depth, gt=gt, gmap=gmap, pts=pts, nrm=nrm, op=op, on=on, sensor=sensor, pin=pin)
    gmap.close()


@pytest.mark.parametrize("level", [0, 1, 2])
@pytest.mark.parametrize("compensated", [False, True])
def test_icp_counters_exact_at_common_pose(frame_and_model, level, compensated):
    """With one iteration at one level, the reduction both sides return is that of a single pass at
    the start pose: trackKernel's result codes are the same pixel for pixel, so the inlier and
    failure counters agree exactly."""
    s = frame_and_model
    se = s["se"]
    its = [0, 0, 0]
    its[level] = 1
    tracker = se.Tracker(s["gmap"], s["sensor"], se.TrackerConfig(iterations=tuple(its), compensated=compensated))
    P = s["gt"][4].astype(np.float32)
    _, _, sums_g = tracker.track(torch.from_numpy(s["depth"][5]).cuda(), P, s["pts"], s["nrm"])
    _, _, sums_o, iters = O.track(s["depth"][5], s["pin"], P, s["op"], s["on"], cfg=O.tracker_config(iterations=its))
    assert iters == 1
    print(f"level {level}: inliers {sums_g[28]:.0f} dist fails {sums_g[29]:.0f} normal fails {sums_g[30]:.0f}")
    np.testing.assert_array_equal(sums_g[28:], sums_o[28:])
    assert sums_g[28] > 100
    np.testing.assert_allclose(sums_g[:28], sums_o[:28], rtol=2e-4, atol=2e-4 * np.abs(sums_o[:28]).max())
    tracker.close()


def _oracle_pyramid(depth, levels, pin, neg_y):
    out, d = [], depth
    for l in range(levels):
        if l > 0:
            d = O.half_sample(d)
        p = O.depth_to_points(d, pin, l)
        out.append((d, p, O.points_to_normals(p, neg_y)))
    return out


@pytest.mark.parametrize("w,h,fy_sign,levels", [(640, 480, 1.0, 3), (640, 480, -1.0, 3), (637, 479, 1.0, 4),
                                                 (640, 480, 1.0, 6), (333, 250, -1.0, 1)])
def test_pyramid_levels_bit_exact(frame_and_model, w, h, fy_sign, levels):
    s = frame_and_model
    se = s["se"]
    depth = np.ascontiguousarray(s["depth"][5][:h, :w])
    sensor, pin = _sensor(se, w=w, h=h, fy_sign=fy_sign), _pin(w, h, fy_sign)
    cfg = se.TrackerConfig(iterations=tuple([1] * levels))
    tracker = se.Tracker(s["gmap"], sensor, cfg)
    zeros = torch.zeros((h, w, 3), dtype=torch.float32, device="cuda")
    tracker.track(torch.from_numpy(depth).cuda(), s["gt"][4], zeros, zeros)
    for l, (d, p, n) in enumerate(_oracle_pyramid(depth, levels, pin, fy_sign < 0)):
        gp, gn = tracker.read_pyramid(l)
        print(f"{w}x{h} fy {fy_sign:+.0f} level {l}: {gp.shape[1]}x{gp.shape[0]}, valid normals "
              f"{int(np.any(gn != 0, axis=-1).sum())}")
        assert gp.shape[:2] == d.shape
        np.testing.assert_array_equal(gp[..., 2], np.where(d > 0, d, 0).astype(np.float32))
        np.testing.assert_array_equal(gp, p)
        np.testing.assert_array_equal(gn, n)
    tracker.close()


@pytest.mark.parametrize("iterations", [(10, 5, 4, 3, 2), (6,), (4, 4)])
def test_tracker_level_counts_match_oracle(frame_and_model, iterations):
    s = frame_and_model
    se = s["se"]
    tracker = se.Tracker(s["gmap"], s["sensor"], se.TrackerConfig(iterations=iterations))
    P = s["gt"][4].astype(np.float32)
    ok_g, T_g, sums_g = tracker.track(torch.from_numpy(s["depth"][5]).cuda(), P, s["pts"], s["nrm"])
    ok_o, T_o, sums_o, iters = O.track(s["depth"][5], s["pin"], P, s["op"], s["on"],
                                       cfg=O.tracker_config(iterations=iterations))
    print(iterations, "ok", ok_g, ok_o, "iterations", iters, "pose difference", np.abs(T_g - T_o).max())
    assert ok_g == ok_o
    np.testing.assert_allclose(T_g, T_o, atol=1e-4)
    tracker.close()


def test_tracker_level_count_limits(frame_and_model):
    from supereight2_amd import capi

    s = frame_and_model
    lib = capi.load()
    cfg = s["se"].TrackerConfig().c()
    h = C.c_void_p()
    for n, want in ((0, capi.SE8_ERR_INVALID), (capi.TRACKER_MAX_LEVELS + 1, capi.SE8_ERR_INVALID)):
        cfg.num_levels = n
        assert lib.se8_tracker_create(C.byref(s["sensor"].c()), C.byref(cfg), C.byref(h)) == want
    # a level with no pixels: 40 x 30 has 2 x 1 pixels at level 4 and none at level 5
    small = capi.Pinhole(40, 30, 40.0, 40.0, 19.5, 14.5, 0.4, 6.0)
    cfg.num_levels = 6
    assert lib.se8_tracker_create(C.byref(small), C.byref(cfg), C.byref(h)) == capi.SE8_ERR_INVALID
    assert b"coarsest" in lib.se8_last_error()
    cfg.num_levels = 5
    assert lib.se8_tracker_create(C.byref(small), C.byref(cfg), C.byref(h)) == capi.SE8_OK
    lib.se8_tracker_destroy(h)

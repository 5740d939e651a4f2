"""Device-resident poses: the *_dev entry points give results identical to the host-pose calls.

Workload: the first 6 frames of the seeded tracking spline, run as the reference app's loop
(track against the previous raycast, integrate, raycast) twice: with host poses, and with the pose
kept in device memory from one call to the next (no host round trip).  Tracked poses, allocated
blocks, voxels, octree nodes and the final raycast must be bit-identical; likewise two lidar
street scans integrated with a host pose and with the same pose in device memory.
"""
import numpy as np
import pytest

from supereight2_amd import synth
from tests.helpers import CAM, spline_frames

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

N_FRAMES = 6


def _dumps(m):
    k, ts, t, w = m.dump_blocks()
    nk, nm, nt = m.dump_nodes()
    return k, ts, t.view(np.uint32), w, nk, nm, nt


def test_tracking_loop_device_poses_match_host_poses():
    from supereight2_amd import se

    depth_np, gt = spline_frames(N_FRAMES)
    depth = torch.from_numpy(depth_np).cuda()
    sensor = se.PinholeCamera(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)
    H, W = CAM.height, CAM.width

    def buffers():
        return ([torch.zeros((H, W, 3), device="cuda") for _ in range(2)],
                [torch.zeros((H, W, 3), device="cuda") for _ in range(2)])

    # host poses
    mh = se.TSDFMap(10.24, 0.01)
    ih, th = se.MapIntegrator(mh), se.Tracker(mh, sensor)
    ph, nh = buffers()
    T = gt[0].copy()
    host_poses, host_ok = [], []
    for k in range(N_FRAMES):
        print(f"host-pose frame {k}", flush=True)
        cur, prev = k % 2, (k + 1) % 2
        if k > 0:
            good, T, _ = th.track(depth[k], T, ph[prev], nh[prev])
            host_ok.append(int(good))
        ih.integrateDepth(k, depth[k], sensor, T)
        se.raycast_volume(mh, sensor, T, ph[cur], nh[cur])
        host_poses.append(T.copy())
    torch.cuda.synchronize()

    # device poses: frame k reads pose k - 1 and writes pose k
    md = se.TSDFMap(10.24, 0.01)
    idv, td = se.MapIntegrator(md), se.Tracker(md, sensor)
    pd, nd = buffers()
    T_dev = torch.zeros((N_FRAMES, 16), dtype=torch.float32, device="cuda")
    T_dev[0] = torch.from_numpy(gt[0].reshape(16))
    ok = torch.zeros(N_FRAMES, dtype=torch.int32, device="cuda")
    for k in range(N_FRAMES):
        print(f"device-pose frame {k}", flush=True)
        cur, prev = k % 2, (k + 1) % 2
        if k > 0:
            td.track_device(depth[k], T_dev[k - 1], T_dev[k], pd[prev], nd[prev], ok=ok[k])
        idv.integrateDepth(k, depth[k], sensor, T_dev[k])
        se.raycast_volume(md, sensor, T_dev[k], pd[cur], nd[cur])
    torch.cuda.synchronize()
    dev_poses = T_dev.cpu().numpy().reshape(N_FRAMES, 4, 4)
    assert ok[1:].cpu().numpy().tolist() == host_ok
    np.testing.assert_array_equal(np.stack(host_poses).view(np.uint32), dev_poses.view(np.uint32))
    for a, b in zip(_dumps(mh), _dumps(md)):
        np.testing.assert_array_equal(a, b)
    last = (N_FRAMES - 1) % 2
    assert torch.equal(ph[last], pd[last]) and torch.equal(nh[last], nd[last])
    st_h, st_d = se.stage_times(mh, th), se.stage_times(md, td)
    assert st_d["n_integrations"] == st_h["n_integrations"] == N_FRAMES
    assert st_d["n_tracks"] == st_h["n_tracks"] == N_FRAMES - 1
    for o in (th, td, mh, md):
        o.close()


def test_lidar_device_pose_matches_host_pose():
    from supereight2_amd import se

    spec = synth.LidarSpec()
    poses = synth.street_trajectory(200)[:2].astype(np.float32)
    ranges = synth.render_street_ranges(synth.make_street(0), spec, poses, seed=1).cuda()
    lidar = se.OusterLidar(spec.width, spec.height, spec.near_plane, spec.far_plane, spec.elevations_deg(),
                           spec.azimuths_deg())
    out = []
    for use_dev in (False, True):
        print(f"lidar scans, device pose {use_dev}", flush=True)
        m = se.TSDFMap(204.8, 0.02, truncation_boundary_factor=4.0, max_blocks=1 << 19)
        integ = se.MapIntegrator(m)
        for k in range(2):
            T = torch.from_numpy(poses[k].reshape(16)).cuda() if use_dev else poses[k]
            integ.integrateDepth(k, ranges[k], lidar, T)
        torch.cuda.synchronize()
        out.append(_dumps(m))
        m.close()
    lidar.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)

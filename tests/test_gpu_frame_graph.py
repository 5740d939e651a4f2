"""Frame graphs (se8_frame_graph_*) against the eager device-pose calls, through the C-ABI.

One frame of the tracking loop (track from the previous pose, integrate, raycast) replayed from a
captured HIP graph must give what se8_track_dev + se8_integrate_depth_dev + se8_raycast_dev give
with the same arguments: every pose, check result and raycast bit-identical, and the final maps'
blocks, weights and tsdf bit-identical.  Replays are checked one by one with eager work (queries,
torch operations) and synchronisations between them, and again with the host running ahead.
"""
import numpy as np
import pytest

from tests.helpers import CAM, spline_frames

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

N_FRAMES = 16


def _sensor(se):
    return se.PinholeCamera(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)


def _setup(se, storage, compensated):
    gmap = se.TSDFMap(10.24, 0.01, storage=storage)
    tracker = se.Tracker(gmap, _sensor(se), se.TrackerConfig(compensated=compensated))
    se.set_stage_timing(gmap, tracker, "off")
    return gmap, tracker


@pytest.fixture(scope="module")
def frames():
    depth, gt = spline_frames(N_FRAMES, seed=10)
    return torch.from_numpy(depth).cuda(), gt


@pytest.mark.parametrize("storage,compensated", [("f32", False), ("f16", True)])
@pytest.mark.parametrize("sync_each", [True, False])
def test_graph_replays_equal_eager_calls(frames, storage, compensated, sync_each):
    from supereight2_amd import se

    depth, gt = frames
    sensor = _sensor(se)
    H, W = CAM.height, CAM.width
    shape = (H, W, 3)

    # eager loop
    emap, etr = _setup(se, storage, compensated)
    einteg = se.MapIntegrator(emap)
    e_pts = [torch.zeros(shape, device="cuda") for _ in range(2)]
    e_nrm = [torch.zeros(shape, device="cuda") for _ in range(2)]
    e_T = torch.zeros((N_FRAMES, 16), device="cuda")
    e_ok = torch.zeros((N_FRAMES, 1), dtype=torch.int32, device="cuda")
    e_T[0] = torch.from_numpy(gt[0].reshape(16))
    eager_ray = []
    einteg.integrateDepth(0, depth[0], sensor, e_T[0])
    se.raycast_volume(emap, sensor, e_T[0], e_pts[1], e_nrm[1])
    for k in range(1, N_FRAMES):
        cur, prev = (k - 1) % 2, k % 2
        etr.track_device(depth[k], e_T[k - 1], e_T[k], e_pts[prev], e_nrm[prev], ok=e_ok[k])
        einteg.integrateDepth(k, depth[k], sensor, e_T[k])
        se.raycast_volume(emap, sensor, e_T[k], e_pts[cur], e_nrm[cur])
        eager_ray.append((e_pts[cur].cpu().numpy().copy(), e_nrm[cur].cpu().numpy().copy()))
    torch.cuda.synchronize()

    # graph loop: frame 0 eagerly into pair 1, then one replay per frame
    gmap, gtr = _setup(se, storage, compensated)
    ginteg = se.MapIntegrator(gmap)
    g_pts = [torch.zeros(shape, device="cuda") for _ in range(2)]
    g_nrm = [torch.zeros(shape, device="cuda") for _ in range(2)]
    g_T = torch.zeros((N_FRAMES, 16), device="cuda")
    g_ok = torch.zeros((N_FRAMES, 1), dtype=torch.int32, device="cuda")
    g_T[0] = torch.from_numpy(gt[0].reshape(16))
    ginteg.integrateDepth(0, depth[0], sensor, g_T[0])
    se.raycast_volume(gmap, sensor, g_T[0], g_pts[1], g_nrm[1])
    graph = se.FrameGraph(gtr, gmap, sensor, g_pts, g_nrm)
    probe = torch.from_numpy(gt[0][:3, 3].reshape(1, 3).astype(np.float32)).cuda()
    for k in range(1, N_FRAMES):
        graph.launch(k, depth[k], g_T[k - 1], g_T[k], ok=g_ok[k])
        if sync_each:
            torch.cuda.synchronize()
            cur = (k - 1) % 2
            gp, gn = g_pts[cur].cpu().numpy(), g_nrm[cur].cpu().numpy()
            ep, en = eager_ray[k - 1]
            print(f"frame {k}: pose equal {torch.equal(g_T[k], e_T[k])}, raycast equal "
                  f"{np.array_equal(gp, ep) and np.array_equal(gn, en)}", flush=True)
            np.testing.assert_array_equal(g_T[k].cpu().numpy(), e_T[k].cpu().numpy())
            np.testing.assert_array_equal(gp, ep)
            np.testing.assert_array_equal(gn, en)
            # eager work on the same map and device between replays
            gmap.getData(probe.cpu().numpy())
            _ = (g_pts[cur] * 2.0).sum().item()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(g_T.cpu().numpy(), e_T.cpu().numpy())
    np.testing.assert_array_equal(g_ok.cpu().numpy(), e_ok.cpu().numpy())
    last = (N_FRAMES - 2) % 2
    np.testing.assert_array_equal(g_pts[last].cpu().numpy(), eager_ray[-1][0])
    np.testing.assert_array_equal(g_nrm[last].cpu().numpy(), eager_ray[-1][1])
    gk, gts, gtsdf, gw = gmap.dump_blocks()
    ek, ets, etsdf, ew = emap.dump_blocks()
    np.testing.assert_array_equal(gk, ek)
    np.testing.assert_array_equal(gts, ets)
    np.testing.assert_array_equal(gw, ew)
    np.testing.assert_array_equal(gtsdf, etsdf)
    assert int(g_ok[1:].sum().item()) == N_FRAMES - 1
    print(f"{storage} compensated={compensated} sync_each={sync_each}: {N_FRAMES - 1} replays equal, "
          f"{len(gk)} blocks")
    graph.close()
    for o in (etr, emap, gtr, gmap):
        o.close()


def test_graph_refuses_timed_and_multires_maps():
    from supereight2_amd import capi, se

    sensor = _sensor(se)
    shape = (CAM.height, CAM.width, 3)
    pts = [torch.zeros(shape, device="cuda") for _ in range(2)]
    nrm = [torch.zeros(shape, device="cuda") for _ in range(2)]
    gmap = se.TSDFMap(10.24, 0.01)
    tracker = se.Tracker(gmap, sensor)
    se.set_stage_timing(gmap, tracker, "stages")
    with pytest.raises(capi.Se8Error, match="stage timing"):
        se.FrameGraph(tracker, gmap, sensor, pts, nrm)
    se.set_stage_timing(gmap, tracker, "off")
    graph = se.FrameGraph(tracker, gmap, sensor, pts, nrm)
    se.set_stage_timing(gmap, tracker, "stages")
    T = torch.eye(4, device="cuda").reshape(16).contiguous()
    with pytest.raises(capi.Se8Error, match="stage timing"):
        graph.launch(1, torch.zeros((CAM.height, CAM.width), device="cuda"), T, T)
    graph.close()
    tracker.close()
    gmap.close()
    mr = se.TSDFMap(10.24, 0.01, storage="multires")
    tr2 = se.Tracker(mr, sensor)
    with pytest.raises(capi.Se8Error):
        se.FrameGraph(tr2, mr, sensor, pts, nrm)
    tr2.close()
    mr.close()

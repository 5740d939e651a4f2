"""Stage-timing modes (se8_set_stage_timing) on the GPU.

The same tracking loop (device-resident poses) run with timing off, per stage and per pyramid level
gives bit-identical poses, blocks and voxels; timing off records no times and leaves the host free
to run ahead (no call waits); the level mode's per-level times cover every Gauss-Newton launch; a
capacity overflow during untimed integrations is reported by se8_map_get_info and by the first
integration after timing is switched back on.
"""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import CAM, spline_frames

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

N_FRAMES = 6


def _sensor(se):
    return se.PinholeCamera(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)


def _loop(se, mode):
    depth_np, gt = spline_frames(N_FRAMES)
    depth = torch.from_numpy(depth_np).cuda()
    sensor = _sensor(se)
    H, W = CAM.height, CAM.width
    m = se.TSDFMap(10.24, 0.01)
    integ, tr = se.MapIntegrator(m), se.Tracker(m, sensor)
    se.set_stage_timing(m, tr, mode)
    pts = [torch.zeros((H, W, 3), device="cuda") for _ in range(2)]
    nrm = [torch.zeros((H, W, 3), device="cuda") for _ in range(2)]
    T_dev = torch.zeros((N_FRAMES, 16), dtype=torch.float32, device="cuda")
    T_dev[0] = torch.from_numpy(gt[0].reshape(16))
    for k in range(N_FRAMES):
        cur, prev = k % 2, (k + 1) % 2
        if k > 0:
            tr.track_device(depth[k], T_dev[k - 1], T_dev[k], pts[prev], nrm[prev])
        integ.integrateDepth(k, depth[k], sensor, T_dev[k])
        se.raycast_volume(m, sensor, T_dev[k], pts[cur], nrm[cur])
    torch.cuda.synchronize()
    st = se.stage_times(m, tr)
    keys, ts, tsdf, w = m.dump_blocks()
    out = (T_dev.cpu().numpy(), keys, ts, tsdf.view(np.uint32), w, pts[(N_FRAMES - 1) % 2].cpu().numpy())
    tr.close()
    m.close()
    return out, st


def test_timing_modes_give_identical_results():
    from supereight2_amd import se

    res = {}
    for mode in ("off", "stages", "levels"):
        print(f"timing {mode}", flush=True)
        res[mode] = _loop(se, mode)
    for mode in ("stages", "levels"):
        for a, b in zip(res["off"][0], res[mode][0]):
            np.testing.assert_array_equal(a, b)
    st_off, st_st, st_lv = res["off"][1], res["stages"][1], res["levels"][1]
    assert st_off["n_integrations"] == st_off["n_raycasts"] == st_off["n_tracks"] == 0
    assert st_st["n_integrations"] == N_FRAMES and st_st["n_tracks"] == N_FRAMES - 1
    assert st_st["sum_update"] > 0 and st_st["sum_raycast"] > 0 and st_st["sum_tracking"] > 0
    assert not any(st_st["sum_track_level"])
    # levels: every Gauss-Newton launch is inside one level's interval, which lies inside the track
    assert st_lv["n_track_level_launches"][:3] == [10 * (N_FRAMES - 1), 5 * (N_FRAMES - 1), 4 * (N_FRAMES - 1)]
    assert all(x > 0 for x in st_lv["sum_track_level"][:3])
    print("level ms", st_lv["sum_track_level"][:3], "tracking ms", st_lv["sum_tracking"])
    assert sum(st_lv["sum_track_level"]) < st_lv["sum_tracking"]


def test_overflow_without_timing_is_reported():
    from supereight2_amd import capi, se

    depth, poses = spline_frames(2)
    sensor = _sensor(se)
    cap = 512
    gmap = se.TSDFMap(10.24, 0.01, max_blocks=cap)
    se.set_stage_timing(gmap, None, "off")
    integ = se.MapIntegrator(gmap)
    integ.integrateDepth(0, torch.from_numpy(depth[0]).cuda(), sensor, poses[0])  # needs ~25k blocks
    integ.integrateDepth(1, torch.from_numpy(depth[1]).cuda(), sensor, poses[1])  # no host check: accepted
    torch.cuda.synchronize()
    info = capi.MapInfo()
    assert gmap.lib.se8_map_get_info(gmap.h, C.byref(info)) == capi.SE8_ERR_CAPACITY
    assert info.num_blocks == cap
    se.set_stage_timing(gmap, None, "stages")
    with pytest.raises(capi.Se8Error, match="capacity"):
        integ.integrateDepth(2, torch.from_numpy(depth[1]).cuda(), sensor, poses[1])
    keys, _, _, _ = gmap.dump_blocks(with_data=False)
    assert len(keys) == cap
    gmap.close()

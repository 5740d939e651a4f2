"""se8_raycast_dev_prepare_track and se8_track_dev_prepare_integrate against the per-call loop,
through the C-ABI.

Raycasting frame k with the pyramid of frame k+1 in the same launch, tracking frame k+1 (which
skips its own pyramid) with the start of its integration in the tracker's last launch, then
integrating it (which skips its frame start), must give what the separate calls give: every pose,
check result and raycast bit-identical, and the final maps' blocks, weights and tsdf
bit-identical.  A track call whose depth or start pose differs from the prepared ones runs its own
pyramid.
"""
import numpy as np
import pytest

from tests.helpers import CAM, spline_frames

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

N_FRAMES = 16


def _sensor(se):
    return se.PinholeCamera(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)


@pytest.fixture(scope="module")
def frames():
    depth, gt = spline_frames(N_FRAMES, seed=10)
    return torch.from_numpy(depth).cuda(), gt


def _run(se, depth, gt, storage, compensated, fused, swap_at=None):
    # fused: each raycast prepares the next track's pyramid and each track the next integration
    sensor = _sensor(se)
    gmap = se.TSDFMap(10.24, 0.01, storage=storage)
    tracker = se.Tracker(gmap, sensor, se.TrackerConfig(compensated=compensated))
    se.set_stage_timing(gmap, tracker, "off")
    integ = se.MapIntegrator(gmap)
    shape = (CAM.height, CAM.width, 3)
    pts = [torch.zeros(shape, device="cuda") for _ in range(2)]
    nrm = [torch.zeros(shape, device="cuda") for _ in range(2)]
    T = torch.zeros((N_FRAMES, 16), device="cuda")
    ok = torch.zeros((N_FRAMES, 1), dtype=torch.int32, device="cuda")
    T[0] = torch.from_numpy(gt[0].reshape(16))
    rays = []
    for k in range(N_FRAMES):
        cur, prev = k % 2, (k + 1) % 2
        if k > 0:
            # at swap_at the prepared depth is replaced by a copy: the track must run its own pyramid
            d = depth[k].clone() if k == swap_at else depth[k]
            if fused:
                se.track_device_prepare_integrate(tracker, d, T[k - 1], T[k], pts[prev], nrm[prev], gmap, ok=ok[k])
            else:
                tracker.track_device(d, T[k - 1], T[k], pts[prev], nrm[prev], ok=ok[k])
        integ.integrateDepth(k, depth[k], sensor, T[k])
        if fused and k + 1 < N_FRAMES:
            se.raycast_volume_prepare_track(gmap, sensor, T[k], pts[cur], nrm[cur], tracker, depth[k + 1])
        else:
            se.raycast_volume(gmap, sensor, T[k], pts[cur], nrm[cur])
        rays.append((pts[cur].cpu().numpy().copy(), nrm[cur].cpu().numpy().copy()))
    torch.cuda.synchronize()
    out = dict(T=T.cpu().numpy(), ok=ok.cpu().numpy(), rays=rays, blocks=gmap.dump_blocks())
    tracker.close()
    gmap.close()
    return out


@pytest.mark.parametrize("storage,compensated", [("f32", False), ("f16", True)])
def test_prepared_pyramid_equals_per_call_loop(frames, storage, compensated):
    from supereight2_amd import se

    depth, gt = frames
    ref = _run(se, depth, gt, storage, compensated, fused=False)
    for swap_at in (None, 7):
        got = _run(se, depth, gt, storage, compensated, fused=True, swap_at=swap_at)
        np.testing.assert_array_equal(got["T"], ref["T"])
        np.testing.assert_array_equal(got["ok"], ref["ok"])
        for k, ((gp, gn), (rp, rn)) in enumerate(zip(got["rays"], ref["rays"])):
            assert np.array_equal(gp, rp) and np.array_equal(gn, rn), f"raycast of frame {k} differs"
        for a, b in zip(got["blocks"], ref["blocks"]):
            np.testing.assert_array_equal(a, b)
        assert int(got["ok"][1:].sum()) == N_FRAMES - 1
        print(f"{storage} compensated={compensated} swap_at={swap_at}: {N_FRAMES} frames equal, "
              f"{len(got['blocks'][0])} blocks", flush=True)

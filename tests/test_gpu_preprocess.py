"""Input side on the GPU: se8_downsample_depth(_u16) against the oracle, and the reference's 3-frame
ICL-NUIM sequence run end to end (reader -> fused uint16 conversion and median downsampling ->
integration, raycast and tracking) against the oracle on the same frames.

Downsampling: output depth bit-exact and the index map equal (equal depths resolved by the region's
row-major order in both) for ratios 1, 2, 4 (register path) and 3, 5, 8 (general path) on images
with many equal depths, NaN, zeros and negatives; the fused uint16 conversions (RAW millimetres
divided by 1000, TUM values times 1/5000); bad sizes refused.

End to end (test/system/multires_ofusion/config.yaml: 640x480 at fx 481.2, fy -480 (a left-handed
image frame), cx 319.5, cy 239.5, downsampling factor 2, 10.24 m map at 1 cm): with ground-truth
poses, allocated blocks, node keys, voxels and the raycast bit-exact after each frame; tracking
frames 2 and 3 from the previous ground-truth pose, the tracked pose within 1e-4 of the oracle's and
within 5 mm / 0.3 degrees of the ground truth.
"""
import numpy as np
import pytest

from oracle import oracle as O
from supereight2_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FIX = __file__.rsplit("/", 1)[0] + "/golden/tum_icl3"


def _random_depth(rng, h, w):
    d = rng.integers(0, 8, (h, w)).astype(np.float32) * np.float32(0.25)
    d[rng.random((h, w)) < 0.08] = np.nan
    d[rng.random((h, w)) < 0.04] = -0.5
    return d


@pytest.mark.parametrize("shape,r", [((480, 640), 1), ((480, 640), 2), ((480, 640), 4), ((360, 480), 3),
                                     ((480, 640), 5), ((480, 640), 8)])
def test_downsample_matches_oracle(shape, r):
    from supereight2_amd import se

    rng = np.random.default_rng(r)
    d = _random_depth(rng, *shape)
    h, w = shape[0] // r, shape[1] // r
    print(f"downsample {shape} ratio {r}", flush=True)
    out = torch.full((h, w), -7.0, dtype=torch.float32, device="cuda")
    mp = torch.zeros((h, w), dtype=torch.int64, device="cuda")
    se.downsample_depth(torch.from_numpy(d).cuda(), out, mp)
    eo, em = O.downsample_depth(d, w, h)
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), eo.view(np.uint32))
    np.testing.assert_array_equal(mp.cpu().numpy().astype(np.uint64), em)


@pytest.mark.parametrize("divide,scale", [(1, 1000.0), (0, 1.0 / 5000.0)])
def test_downsample_u16_matches_oracle(divide, scale):
    from supereight2_amd import se

    rng = np.random.default_rng(7 + divide)
    raw = rng.integers(0, 20000, (480, 640)).astype(np.uint16)
    raw[rng.random(raw.shape) < 0.1] = 0
    raw_dev = torch.from_numpy(raw.view(np.int16)).cuda()
    for r in (1, 2, 8):
        h, w = 480 // r, 640 // r
        out = torch.zeros((h, w), dtype=torch.float32, device="cuda")
        mp = torch.zeros((h, w), dtype=torch.int64, device="cuda")
        se.downsample_depth_u16(raw_dev, np.float32(scale), divide, out, mp)
        conv = O.depth_from_u16(raw, np.float32(scale), divide)
        eo, em = O.downsample_depth(conv, w, h)
        np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), eo.view(np.uint32))
        np.testing.assert_array_equal(mp.cpu().numpy().astype(np.uint64), em)


def test_downsample_rejects_bad_sizes():
    from supereight2_amd import capi, se

    d = torch.zeros((480, 640), dtype=torch.float32, device="cuda")
    for h, w in ((240, 300), (160, 320), (7, 9)):
        with pytest.raises(capi.Se8Error, match="integer multiple|bad image sizes"):
            se.downsample_depth(d, torch.zeros((h, w), dtype=torch.float32, device="cuda"))


def _compare_maps(gmap, omap):
    gk, gts, gt, gw = gmap.dump_blocks()
    ok, ots, ot, ow = omap.dump_blocks()
    np.testing.assert_array_equal(gk, ok)
    np.testing.assert_array_equal(gts, ots)
    np.testing.assert_array_equal(gw, ow)
    np.testing.assert_array_equal(gt.view(np.uint32), ot.view(np.uint32))
    for a, b in zip(gmap.dump_nodes(), omap.dump_nodes()):
        np.testing.assert_array_equal(a, b)
    return len(gk)


def _rot_angle_deg(Ra, Rb):
    return synth.rotation_angle_deg(np.asarray(Ra, np.float64).T @ np.asarray(Rb, np.float64))


def test_icl_sequence_end_to_end():
    from supereight2_amd import se

    reader = se.Reader("tum", FIX, FIX + "/groundtruth.txt")
    assert reader.num_frames == 3
    dsf = 2
    W, H = 640 // dsf, 480 // dsf
    fx, fy = np.float32(481.2) / np.float32(dsf), np.float32(-480.0) / np.float32(dsf)
    cx = (np.float32(319.5) + np.float32(0.5)) / np.float32(dsf) - np.float32(0.5)
    cy = (np.float32(239.5) + np.float32(0.5)) / np.float32(dsf) - np.float32(0.5)
    sensor = se.PinholeCamera(W, H, float(fx), float(fy), float(cx), float(cy), 0.4, 6.0)
    pin = O.pinhole(W, H, float(fx), float(fy), float(cx), float(cy), 0.4, 6.0)
    gmap, omap = se.TSDFMap(10.24, 0.01), O.OracleMap(10.24, 0.01)
    integ = se.MapIntegrator(gmap)
    tracker = se.Tracker(gmap, sensor)
    pts = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    nrm = torch.zeros_like(pts)
    o_pts = o_nrm = None
    for k in range(3):
        st, raw, scale, divide, T_gt = reader.nextDataRaw()
        assert st == se.ReaderStatus.ok
        print(f"ICL frame {k}", flush=True)
        depth = torch.zeros((H, W), dtype=torch.float32, device="cuda")
        se.downsample_depth_u16(torch.from_numpy(raw.view(np.int16)).cuda(), scale, divide, depth)
        o_depth, _ = O.downsample_depth(O.depth_from_u16(raw, scale, divide), W, H)
        g_depth = depth.cpu().numpy()
        np.testing.assert_array_equal(g_depth.view(np.uint32), o_depth.view(np.uint32))
        assert (g_depth > 0).mean() > 0.9
        if k > 0:
            # track from the previous frame's ground-truth pose against its raycast
            ok_g, T_g, _ = tracker.track(depth, T_prev, pts, nrm)
            ok_o, T_o, _, _ = O.track(o_depth, pin, T_prev, o_pts, o_nrm)
            print(f"  track ok {ok_g}/{ok_o}, GT distance {np.linalg.norm(T_g[:3, 3] - T_gt[:3, 3]) * 1000:.2f} mm, "
                  f"{_rot_angle_deg(T_g[:3, :3], T_gt[:3, :3]):.3f} deg", flush=True)
            assert ok_g and ok_o
            np.testing.assert_allclose(T_g, T_o, atol=1e-4)
            assert np.linalg.norm(T_g[:3, 3] - T_gt[:3, 3]) < 5e-3
            assert _rot_angle_deg(T_g[:3, :3], T_gt[:3, :3]) < 0.3
        integ.integrateDepth(k, depth, sensor, T_gt)
        omap.integrate(k, o_depth, pin, T_gt)
        se.raycast_volume(gmap, sensor, T_gt, pts, nrm)
        o_pts, o_nrm, _ = omap.raycast(pin, T_gt)
        torch.cuda.synchronize()
        n = _compare_maps(gmap, omap)
        np.testing.assert_array_equal(pts.cpu().numpy().reshape(o_pts.shape), o_pts)
        np.testing.assert_array_equal(nrm.cpu().numpy().reshape(o_nrm.shape), o_nrm)
        hit = np.abs(o_pts).sum(-1) > 0
        print(f"  blocks {n}, raycast hits {hit.mean():.3f}", flush=True)
        assert hit.mean() > 0.8
        T_prev = T_gt
    tracker.close()
    gmap.close()
    reader.close()

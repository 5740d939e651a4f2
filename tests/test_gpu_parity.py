"""GPU parity of the HIP path (through the C-ABI) against the CPU oracle.

Bar: block allocation (key set, node set, child masks, timestamps) and the FP32 TSDF voxels are
bit-exact; raycast vertex/normal maps are bit-exact; ICP poses agree to 1e-4 (the reference's
reduction order is unspecified, so sums differ in the last bits).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import CAM, circle_frames, spline_frames

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _se():
    from supereight2_amd import se

    return se


def _sensor(se):
    return se.PinholeCamera(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)


def _pin():
    return O.pinhole(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)


@pytest.fixture(scope="module")
def integrated():
    """GPU map and oracle map after integrating the same 4 circle frames."""
    se = _se()
    depth, poses = circle_frames(50, 4)
    sensor = _sensor(se)
    gmap = se.TSDFMap(10.24, 0.01)
    integ = se.MapIntegrator(gmap)
    omap = O.OracleMap(10.24, 0.01)
    pin = _pin()
    stats = []
    for k in range(len(depth)):
        print(f"integrate frame {k}", flush=True)
        d = torch.from_numpy(depth[k]).cuda()
        integ.integrateDepth(k, d, sensor, poses[k])
        torch.cuda.synchronize()
        stats.append(omap.integrate(k, depth[k], pin, poses[k]))
    return se, gmap, omap, depth, poses, stats


def test_allocation_block_set_bit_exact(integrated):
    se, gmap, omap, _, _, stats = integrated
    gk, gts, _, _ = gmap.dump_blocks(with_data=False)
    ok, ots, _, _ = omap.dump_blocks(with_data=False)
    print("blocks gpu/oracle", len(gk), len(ok), "oracle stats", stats)
    assert len(gk) == len(ok)
    np.testing.assert_array_equal(gk, ok)
    np.testing.assert_array_equal(gts, ots)


def test_octree_topology_bit_exact(integrated):
    se, gmap, omap, _, _, _ = integrated
    gk, gm, gt = gmap.dump_nodes()
    ok, om, ot = omap.dump_nodes()
    assert len(gk) == len(ok)
    np.testing.assert_array_equal(gk, ok)
    np.testing.assert_array_equal(gm, om)
    np.testing.assert_array_equal(gt, ot)


def test_tsdf_voxels_bit_exact(integrated):
    se, gmap, omap, _, _, _ = integrated
    gk, _, gt, gw = gmap.dump_blocks()
    ok, _, ot, ow = omap.dump_blocks()
    np.testing.assert_array_equal(gk, ok)
    diff = np.abs(gt - ot)
    print("max |dtsdf|", diff.max(), "n differing voxels", int((gt != ot).sum()))
    np.testing.assert_array_equal(gw, ow)
    np.testing.assert_array_equal(gt, ot)


def test_update_list_matches_oracle(integrated):
    """integrateDepth's updated_octants (map_integrator_impl.hpp:77-80): the last integration's
    frustum + new blocks, as a key set, on the dense-table (claim) allocation path."""
    se, gmap, omap, _, _, stats = integrated
    g, o = gmap.update_list(), omap.update_list()
    print("updated blocks gpu/oracle", len(g), len(o), "oracle frustum + new", stats[-1]["frustum"] + stats[-1]["new"])
    assert len(np.unique(g)) == len(g)
    np.testing.assert_array_equal(g, o)
    assert len(o) == stats[-1]["frustum"] + stats[-1]["new"]


@pytest.mark.parametrize("dim", [10.24, 81.92])
def test_update_list_every_frame(dim):
    """The update list after each of 3 frames, on the claim path (10.24 m: dense block table) and the
    sort path (81.92 m: brick index)."""
    se = _se()
    depth, poses = circle_frames(50, 3)
    sensor, pin = _sensor(se), _pin()
    gmap = se.TSDFMap(dim, 0.01)
    integ = se.MapIntegrator(gmap)
    omap = O.OracleMap(dim, 0.01)
    for k in range(len(depth)):
        print(f"map {dim} m: integrate frame {k}", flush=True)
        integ.integrateDepth(k, torch.from_numpy(depth[k]).cuda(), sensor, poses[k])
        st = omap.integrate(k, depth[k], pin, poses[k])
        g, o = gmap.update_list(), omap.update_list()
        assert len(o) == st["frustum"] + st["new"]
        np.testing.assert_array_equal(g, o)
    gmap.close()


def test_aabb_matches(integrated):
    se, gmap, omap, _, _, _ = integrated
    gv, gmin, gmax = gmap.aabb()
    ov, omin, omax = omap.aabb()
    assert gv and ov
    np.testing.assert_array_equal(gmin, omin)
    np.testing.assert_array_equal(gmax, omax)


def test_raycast_bit_exact(integrated):
    se, gmap, omap, _, poses, _ = integrated
    sensor = _sensor(se)
    T = poses[-1]
    pts = torch.zeros((CAM.height, CAM.width, 3), dtype=torch.float32, device="cuda")
    nrm = torch.zeros_like(pts)
    se.raycast_volume(gmap, sensor, T, pts, nrm)
    torch.cuda.synchronize()
    gp, gn = pts.cpu().numpy(), nrm.cpu().numpy()
    op, on, _ = omap.raycast(_pin(), T)
    hit = np.any(op != 0, axis=-1)
    print("hit fraction", hit.mean(), "point mismatches", int(np.any(gp != op, axis=-1).sum()),
          "normal mismatches", int(np.any(gn != on, axis=-1).sum()))
    np.testing.assert_array_equal(gp, op)
    np.testing.assert_array_equal(gn, on)
    assert hit.mean() > 0.8


def test_field_queries_match_oracle(integrated):
    se, gmap, omap, _, poses, _ = integrated
    rng = np.random.default_rng(5)
    op, _, _ = omap.raycast(_pin(), poses[-1])
    hits = op[np.any(op != 0, axis=-1)]
    pts = hits[rng.choice(len(hits), 2000, replace=False)] + rng.normal(0, 0.02, (2000, 3)).astype(np.float32)
    gi, gv = gmap.getFieldInterp(pts)
    gg, ggv = gmap.getFieldGrad(pts)
    gt, gw = gmap.getData(pts)
    for i, p in enumerate(pts):
        oi = omap.field_interp(p)
        assert (oi is not None) == bool(gv[i])
        if oi is not None:
            assert np.float32(oi) == gi[i]
        og = omap.field_grad(p)
        assert (og is not None) == bool(ggv[i])
        if og is not None:
            np.testing.assert_array_equal(og.astype(np.float32), gg[i])
        ot, ow = omap.get_data(p)
        assert np.float32(ot) == gt[i] and np.float32(ow) == gw[i]


def test_track_matches_oracle():
    se = _se()
    depth, poses = spline_frames(6)
    sensor = _sensor(se)
    gmap = se.TSDFMap(10.24, 0.01)
    integ = se.MapIntegrator(gmap)
    omap = O.OracleMap(10.24, 0.01)
    pin = _pin()
    # map the first 5 frames at ground truth, then track frame 5 from frame 4's pose
    for k in range(5):
        integ.integrateDepth(k, torch.from_numpy(depth[k]).cuda(), sensor, poses[k])
        omap.integrate(k, depth[k], pin, poses[k])
    pts = torch.zeros((CAM.height, CAM.width, 3), dtype=torch.float32, device="cuda")
    nrm = torch.zeros_like(pts)
    se.raycast_volume(gmap, sensor, poses[4], pts, nrm)
    op, on, _ = omap.raycast(pin, poses[4])
    np.testing.assert_array_equal(pts.cpu().numpy(), op)
    for compensated in (False, True):
        tracker = se.Tracker(gmap, sensor, se.TrackerConfig(compensated=compensated))
        ok_g, T_g, sums_g = tracker.track(torch.from_numpy(depth[5]).cuda(), poses[4], pts, nrm)
        ok_o, T_o, sums_o, iters = O.track(depth[5], pin, poses[4], op, on)
        print("compensated", compensated, "ok", ok_g, ok_o, "iters", iters)
        print("pose gpu\n", T_g, "\npose oracle\n", T_o, "\ngt\n", poses[5])
        assert ok_g == ok_o
        np.testing.assert_allclose(T_g, T_o, atol=1e-4)
        # inlier and failure counts: poses differ in the last bits (sin/cos and summation order),
        # which flips a few pixels at the association gates
        np.testing.assert_allclose(sums_g[28:], sums_o[28:], rtol=1e-3, atol=10)
        # tracking recovers the ground-truth motion to within a few mm
        assert np.abs(T_g[:3, 3] - poses[5][:3, 3]).max() < 0.01
        tracker.close()


def test_map_interpolation_kat():
    """test/map/map/map_unittest.cpp:127-208 through the C-ABI."""
    se = _se()
    m = se.TSDFMap(32.0, 1.0)
    bs = 8
    block_coords = [(0, 0, 0), (bs, 0, 0), (0, bs, 0), (bs, bs, 0), (0, 0, bs), (bs, 0, bs), (0, bs, bs), (bs, bs, bs)]
    for i, bc in enumerate(block_coords):
        got = m.allocate_blocks([bc])
        np.testing.assert_array_equal(got[0], bc)
        m.write_block(bc, np.full(512, float(i), np.float32), np.ones(512, np.float32))
    queries = [(-12, -12, -12), (-8, -12, -12), (-12, -8, -12), (-8, -8, -12), (-12, -12, -8), (-8, -12, -8),
               (-12, -8, -8), (-8, -8, -8), (2, 2, 2)]
    want = [0, 0.5, 1, 1.5, 2, 2.5, 3, 3.5, None]
    vals, valid = m.getFieldInterp(np.array(queries, np.float32))
    for v, ok, w in zip(vals, valid, want):
        if w is None:
            assert not ok
        else:
            assert ok and v == w


def test_allocation_kats():
    """test/map/octree_allocator/octree_allocator_unittest.cpp:280-378 through the C-ABI."""
    se = _se()
    m = se.TSDFMap(256.0, 1.0)
    vox = np.array([[33, 104, 55], [113, 144, 155], [233, 44, 255]], np.int32)
    got = m.allocate_blocks(vox)
    want = sorted([tuple(int(x) for x in (v >> 3) << 3) for v in vox], key=lambda c: O.encode_code(c))
    np.testing.assert_array_equal(got, np.array(want))
    # ancestors of every block at scales 4..8 exist with the right coordinates
    keys, masks, _ = m.dump_nodes()
    node_set = set(int(k) for k in keys)
    for v in vox:
        for s in range(4, 9):
            k, _ = O.encode_key(((v >> s) << s), s)
            assert k in node_set
    vox2 = np.concatenate([vox, [[0, 0, 0]]]).astype(np.int32)
    new = m.allocate_blocks(vox2, only_new=True)
    np.testing.assert_array_equal(new, [[0, 0, 0]])
    assert len(m.allocate_blocks(vox2, only_new=True)) == 0
    assert len(m.allocate_blocks(vox2)) == 4

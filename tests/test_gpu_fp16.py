"""FP16-TSDF storage on the GPU against the oracle with FP16 storage (every updated tsdf rounded to the
nearest binary16, as the GPU's packed store does), through the C-ABI.

Bar: block set, weights and tsdf values bit-exact; the raycast's vertex and normal maps bit-exact on
both block tables (the dense table of a 10.24 m map: k_raycast<3>; the brick index of an 81.92 m
map: k_raycast<1>); getData / getFieldInterp / getFieldGrad bit-exact (k_query<3>, k_query<1>).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import CAM, circle_frames

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _sensor(se):
    return se.PinholeCamera(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)


def _pin():
    return O.pinhole(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)


def _integrate(dim, n):
    from supereight2_amd import se

    depth, poses = circle_frames(50, n)
    sensor = _sensor(se)
    gmap = se.TSDFMap(dim, 0.01, storage="f16")
    integ = se.MapIntegrator(gmap)
    omap = O.OracleMap(dim, 0.01, storage="f16")
    pin = _pin()
    for k in range(n):
        print(f"integrate frame {k} (map {dim} m)", flush=True)
        integ.integrateDepth(k, torch.from_numpy(depth[k]).cuda(), sensor, poses[k])
        omap.integrate(k, depth[k], pin, poses[k])
    torch.cuda.synchronize()
    return se, gmap, omap, poses


@pytest.fixture(scope="module")
def dense():
    se, gmap, omap, poses = _integrate(10.24, 4)
    assert gmap.octree_size() == 1024
    yield se, gmap, omap, poses
    gmap.close()


@pytest.fixture(scope="module")
def brick():
    se, gmap, omap, poses = _integrate(81.92, 2)
    assert gmap.octree_size() == 8192  # above the dense table's limit: lookups use the brick index
    yield se, gmap, omap, poses
    gmap.close()


def _check_voxels(gmap, omap):
    gk, gts, gt, gw = gmap.dump_blocks()
    ok, ots, ot, ow = omap.dump_blocks()
    print("blocks", len(gk), "tsdf values not on the half grid:", int((O.half_round(gt[:64]) != gt[:64]).sum()))
    np.testing.assert_array_equal(gk, ok)
    np.testing.assert_array_equal(gts, ots)
    np.testing.assert_array_equal(gw, ow)
    np.testing.assert_array_equal(gt, ot)
    # stored values are binary16 values
    sample = gt[:: max(1, len(gt) // 200)]
    np.testing.assert_array_equal(O.half_round(sample), sample)


def _check_raycast(se, gmap, omap, T):
    sensor = _sensor(se)
    pts = torch.zeros((CAM.height, CAM.width, 3), dtype=torch.float32, device="cuda")
    nrm = torch.zeros_like(pts)
    se.raycast_volume(gmap, sensor, T, pts, nrm)
    torch.cuda.synchronize()
    gp, gn = pts.cpu().numpy(), nrm.cpu().numpy()
    op, on, _ = omap.raycast(_pin(), T)
    hit = np.any(op != 0, axis=-1)
    print("hit fraction", hit.mean(), "point mismatches", int(np.any(gp != op, axis=-1).sum()),
          "normal mismatches", int(np.any(gn != on, axis=-1).sum()), flush=True)
    np.testing.assert_array_equal(gp, op)
    np.testing.assert_array_equal(gn, on)
    assert hit.mean() > 0.8
    return op


def _check_queries(gmap, omap, op):
    rng = np.random.default_rng(7)
    hits = op[np.any(op != 0, axis=-1)]
    pts = hits[rng.choice(len(hits), 1500, replace=False)] + rng.normal(0, 0.02, (1500, 3)).astype(np.float32)
    gi, gv = gmap.getFieldInterp(pts)
    gg, ggv = gmap.getFieldGrad(pts)
    gt, gw = gmap.getData(pts)
    n_interp = n_grad = 0
    for i, p in enumerate(pts):
        oi = omap.field_interp(p)
        assert (oi is not None) == bool(gv[i])
        if oi is not None:
            assert np.float32(oi) == gi[i]
            n_interp += 1
        og = omap.field_grad(p)
        assert (og is not None) == bool(ggv[i])
        if og is not None:
            np.testing.assert_array_equal(og.astype(np.float32), gg[i])
            n_grad += 1
        ot, ow = omap.get_data(p)
        assert np.float32(ot) == gt[i] and np.float32(ow) == gw[i]
    print("interpolable", n_interp, "with gradient", n_grad)
    assert n_interp > 500 and n_grad > 500


def test_fp16_voxels_bit_exact_dense(dense):
    se, gmap, omap, _ = dense
    _check_voxels(gmap, omap)


def test_fp16_raycast_and_queries_dense_table(dense):
    se, gmap, omap, poses = dense
    op = _check_raycast(se, gmap, omap, poses[-1])
    _check_queries(gmap, omap, op)


def test_fp16_brick_index_map(brick):
    se, gmap, omap, poses = brick
    _check_voxels(gmap, omap)
    op = _check_raycast(se, gmap, omap, poses[-1])
    _check_queries(gmap, omap, op)

"""Block lookups through the two-level brick index.

Maps of at most 4096 voxels a side resolve a block position with one load from a dense table;
larger maps go through the brick id and the brick's slot table.  The other parity tests use
1024-voxel maps, so this one integrates two circle frames into an 81.92 m map at 1 cm (8192 voxels
a side, brick index only) and holds it to the same bar: block set, voxels and the raycast's vertex
and normal maps bit-exact against the oracle.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import CAM, circle_frames

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def test_brick_index_map_matches_oracle():
    from supereight2_amd import se

    depth, poses = circle_frames(50, 2)
    sensor = se.PinholeCamera(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)
    pin = O.pinhole(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)
    gmap, omap = se.TSDFMap(81.92, 0.01), O.OracleMap(81.92, 0.01)
    assert gmap.octree_size() == 8192
    integ = se.MapIntegrator(gmap)
    for k in range(len(depth)):
        print(f"integrate frame {k}", flush=True)
        integ.integrateDepth(k, torch.from_numpy(depth[k]).cuda(), sensor, poses[k])
        omap.integrate(k, depth[k], pin, poses[k])
    torch.cuda.synchronize()
    gk, _, gt, gw = gmap.dump_blocks()
    ok, _, ot, ow = omap.dump_blocks()
    print("blocks gpu/oracle", len(gk), len(ok), flush=True)
    np.testing.assert_array_equal(gk, ok)
    np.testing.assert_array_equal(gw, ow)
    np.testing.assert_array_equal(gt, ot)
    pts = torch.zeros((CAM.height, CAM.width, 3), dtype=torch.float32, device="cuda")
    nrm = torch.zeros_like(pts)
    se.raycast_volume(gmap, sensor, poses[-1], pts, nrm)
    torch.cuda.synchronize()
    op, on, _ = omap.raycast(pin, poses[-1])
    hit = np.any(op != 0, axis=-1)
    print("hit fraction", hit.mean(), flush=True)
    np.testing.assert_array_equal(pts.cpu().numpy(), op)
    np.testing.assert_array_equal(nrm.cpu().numpy(), on)
    assert hit.mean() > 0.8
    gmap.close()

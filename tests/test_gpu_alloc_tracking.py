"""Allocation over small-motion tracking frames is bit-exact against the oracle.

The circle-sequence parity test integrates frames 7.2 degrees apart; tracking frames move at most
2 cm and 1 degree, so after the first frame each frame adds only a few thousand candidate blocks.
Workload: the first 6 frames of the seeded tracking spline integrated at ground truth; block keys,
timestamps, voxels and the octree nodes must equal the oracle's.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import CAM, spline_frames

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

N_FRAMES = 6


def test_tracking_frames_allocation_matches_oracle():
    from supereight2_amd import se

    depth, poses = spline_frames(N_FRAMES)
    sensor = se.PinholeCamera(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)
    pin = O.pinhole(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)
    gmap = se.TSDFMap(10.24, 0.01)
    integ = se.MapIntegrator(gmap)
    omap = O.OracleMap(10.24, 0.01)
    cand = []
    for k in range(N_FRAMES):
        print(f"integrate frame {k}", flush=True)
        integ.integrateDepth(k, torch.from_numpy(depth[k]).cuda(), sensor, poses[k])
        torch.cuda.synchronize()
        cand.append(se.stage_times(gmap)["last_candidates"])
        omap.integrate(k, depth[k], pin, poses[k])
    print("candidate blocks per frame", cand)
    gk, gts, gt, gw = gmap.dump_blocks()
    ok, ots, ot, ow = omap.dump_blocks()
    np.testing.assert_array_equal(gk, ok)
    np.testing.assert_array_equal(gts, ots)
    np.testing.assert_array_equal(gw, ow)
    np.testing.assert_array_equal(gt, ot)
    gnk, gnm, gnt = gmap.dump_nodes()
    onk, onm, ont = omap.dump_nodes()
    np.testing.assert_array_equal(gnk, onk)
    np.testing.assert_array_equal(gnm, onm)
    np.testing.assert_array_equal(gnt, ont)
    gmap.close()

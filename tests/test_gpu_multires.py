"""Multi-resolution TSDF map (SE8_STORAGE_MULTIRES) on the GPU against the oracle.

Five frames 40 apart on the seeded spline, a 10.24 m map at 4 mm (the pixel footprint crosses the
1.5 pixel/voxel threshold inside the room, so blocks move between scales 0 and 1 in both
directions, exercising the hysteresis, the delta down-propagation, the interpolated initialisation
of unobserved children and the up-propagation).  After every frame: block keys, current and
minimum scales, the data and past data of all four scales bit-exact (NaN compared as NaN), node
keys and timestamps equal.  Entry points that read single-resolution voxels refuse the map.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import CAM

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FRAMES = [0, 40, 80, 120, 160]


def _frames():
    from supereight2_amd import synth

    scene = synth.make_scene(10)
    traj = synth.spline_trajectory(300, seed=11)
    rng = np.random.default_rng(3)
    poses = [traj[k].astype(np.float32) for k in FRAMES]
    depth = [synth.render_depth(scene, CAM, T, rng) for T in poses]
    return depth, poses


def _bits(a):
    a = np.ascontiguousarray(a, np.float32).copy()
    a[np.isnan(a)] = np.float32("nan")
    return a.view(np.uint32)


def test_multires_integration_matches_oracle():
    from supereight2_amd import se

    depth, poses = _frames()
    sensor = se.PinholeCamera(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)
    pin = O.pinhole(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)
    gmap = se.TSDFMap(10.24, 0.004, storage="multires", max_blocks=1 << 19)
    omap = O.OracleMap(10.24, 0.004, multires=True)
    integ = se.MapIntegrator(gmap)
    prev = None
    for j in range(len(FRAMES)):
        print(f"multires frame {j}", flush=True)
        integ.integrateDepth(j, torch.from_numpy(depth[j]).cuda(), sensor, poses[j])
        omap.integrate(j, depth[j], pin, poses[j])
        torch.cuda.synchronize()
        gk, gs, gd, gp = gmap.dump_blocks_multires()
        ok, os_, od, op = omap.dump_multires()
        np.testing.assert_array_equal(gk, ok)
        np.testing.assert_array_equal(gs, os_)
        np.testing.assert_array_equal(_bits(gd), _bits(od))
        np.testing.assert_array_equal(_bits(gp), _bits(op))
        for a, b in zip(gmap.dump_nodes(), omap.dump_nodes()):
            np.testing.assert_array_equal(a, b)
        if prev is not None:
            pk, ps = prev
            old = dict(zip(pk.tolist(), ps[:, 0].tolist()))
            moved = [(old[k], s) for k, s in zip(gk.tolist(), gs[:, 0].tolist()) if k in old and old[k] != s]
            print(f"  blocks {len(gk)}, scale changes coarser {sum(a < b for a, b in moved)} "
                  f"finer {sum(a > b for a, b in moved)}", flush=True)
        prev = (gk, gs)
    assert (gs[:, 0] == 0).sum() > 1000 and (gs[:, 0] == 1).sum() > 1000
    q, sc = gmap.structure()
    assert sorted(np.unique(sc.cpu().numpy()).tolist()) == [0, 1]  # faces carry the current scale
    gmap.close()


def test_multires_map_refuses_single_resolution_entry_points():
    from supereight2_amd import capi, se

    m = se.TSDFMap(10.24, 0.01, storage="multires")
    sensor = se.PinholeCamera(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)
    pts = torch.zeros((CAM.height, CAM.width, 3), dtype=torch.float32, device="cuda")
    with pytest.raises(capi.Se8Error, match="multi-resolution"):
        se.raycast_volume(m, sensor, np.eye(4, dtype=np.float32), pts, pts.clone())
    with pytest.raises(capi.Se8Error, match="multi-resolution"):
        m.mesh()
    with pytest.raises(capi.Se8Error, match="multi-resolution"):
        m.getData(np.zeros((1, 3), np.float32))
    with pytest.raises(capi.Se8Error, match="multi-resolution"):
        m.dump_blocks()
    m.close()

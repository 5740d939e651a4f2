"""Spinning-lidar integration (OusterLidar, ouster_lidar.cpp:59-253, ouster_lidar_impl.hpp:15-32)
against the oracle.

Workload: the seeded street (ground plane, 40 box buildings, 200 cylinders, N(0, 2 cm) range
noise) scanned by a 1024 x 128 beam lidar over +-22.5 degrees with a 60 m far plane, into a
204.8 m map at 2 cm voxels (a 16384^3 octree, 14 scales) with a truncation of 4 voxels.  The
device beam model must project bit-exactly like the oracle's; integration must give the oracle's
block keys, timestamps, octree nodes and voxels bit for bit (FP32), and the FP16 map the same
allocation with TSDF values within the fp16 rounding bound of the FP32 map.
"""
import functools

import numpy as np
import pytest

from oracle import oracle as O
from supereight2_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SPEC = synth.LidarSpec()
N_SCANS = 4


def _oracle_lidar(az=None):
    return O.OracleLidar(SPEC.width, SPEC.height, SPEC.near_plane, SPEC.far_plane, SPEC.elevations_deg(),
                         SPEC.azimuths_deg() if az is None else az)


def _gpu_lidar(se, az=None):
    return se.OusterLidar(SPEC.width, SPEC.height, SPEC.near_plane, SPEC.far_plane, SPEC.elevations_deg(),
                          SPEC.azimuths_deg() if az is None else az)


@functools.lru_cache(maxsize=1)
def _street_scans():
    street = synth.make_street(0)
    poses = synth.street_trajectory(200)[:N_SCANS]
    ranges = synth.render_street_ranges(street, SPEC, poses, seed=1).numpy()
    return ranges, poses.astype(np.float32)


def _probe_points():
    rng = np.random.default_rng(3)
    dirs = synth.lidar_directions(SPEC).reshape(-1, 3)
    pts = [dirs * 7.5, dirs * 41.0]
    pts.append(rng.uniform(-70, 70, (20000, 3)))
    # the azimuth seam at +-pi, the poles, the vertical edges of the field of view and the origin
    seam = np.stack([-np.linspace(0.5, 50, 200), np.zeros(200), np.linspace(-5, 5, 200)], 1)
    pts += [seam, seam * [1, -1e-7, 1], np.array([[0, 0, 5.0], [0, 0, -5.0], [0, 0, 0], [1e-30, 0, 0]])]
    e = np.radians(np.linspace(-23.5, 23.5, 400))
    pts.append(np.stack([np.cos(e) * 10, np.zeros_like(e), np.sin(e) * 10], 1))
    return np.concatenate(pts).astype(np.float32)


@pytest.mark.parametrize("offsets", [False, True])
def test_lidar_project_matches_oracle(offsets):
    from supereight2_amd import se

    az = np.linspace(-3.0, 3.0, SPEC.height).astype(np.float32) if offsets else None
    pts = _probe_points()
    print(f"lidar project: {len(pts)} points, azimuth offsets {offsets}", flush=True)
    dev = _gpu_lidar(se, az)
    uv_g, ok_g = dev.project(pts)
    uv_o, ok_o = _oracle_lidar(az).project(pts)
    dev.close()
    np.testing.assert_array_equal(ok_g, ok_o)
    assert ok_g.sum() > len(pts) // 2
    np.testing.assert_array_equal(uv_g[ok_g].view(np.uint32), uv_o[ok_o].view(np.uint32))


def test_lidar_config_rejected():
    from supereight2_amd import capi, se

    with pytest.raises(capi.Se8Error):
        se.OusterLidar(64, 4, 0.5, 60.0, np.array([1.0, 2.0, 0.0, -1.0]))  # not descending
    with pytest.raises(capi.Se8Error):
        se.OusterLidar(64, 4, 2.0, 1.0, np.array([3.0, 1.0, 0.0, -1.0]))  # far < near


def _integrate(se, storage, ranges, poses):
    gmap = se.TSDFMap(204.8, 0.02, truncation_boundary_factor=4.0, storage=storage, max_blocks=1 << 20)
    lidar = _gpu_lidar(se)
    integ = se.MapIntegrator(gmap)
    stats = []
    for k in range(len(ranges)):
        print(f"integrate scan {k} ({storage})", flush=True)
        integ.integrateDepth(k, torch.from_numpy(ranges[k]).cuda(), lidar, poses[k])
        torch.cuda.synchronize()
        st = se.stage_times(gmap)
        stats.append((st["last_frustum_blocks"], st["last_new_blocks"]))
    lidar.close()
    return gmap, stats


def test_lidar_street_integration_matches_oracle():
    from supereight2_amd import se

    ranges, poses = _street_scans()
    gmap, gstats = _integrate(se, "f32", ranges, poses)
    omap = O.OracleMap(204.8, 0.02, mu_factor=4.0)
    olid = _oracle_lidar()
    ostats = []
    for k in range(N_SCANS):
        s = omap.integrate_lidar(k, ranges[k], olid, poses[k])
        ostats.append((s["frustum"], s["new"]))
    print("frustum/new blocks per scan: gpu", gstats, "oracle", ostats)
    assert gstats == ostats
    assert omap.size == 16384 and gmap.octree_size() == 16384
    gnk, gnm, gnt = gmap.dump_nodes()
    onk, onm, ont = omap.dump_nodes()
    np.testing.assert_array_equal(gnk, onk)
    np.testing.assert_array_equal(gnm, onm)
    np.testing.assert_array_equal(gnt, ont)
    gk, gts, gt, gw = gmap.dump_blocks()
    ok, ots, ot, ow = omap.dump_blocks()
    assert len(gk) > 100000
    np.testing.assert_array_equal(gk, ok)
    np.testing.assert_array_equal(gts, ots)
    np.testing.assert_array_equal(gw, ow)
    np.testing.assert_array_equal(gt.view(np.uint32), ot.view(np.uint32))
    gmap.close()


def test_lidar_street_fp16_matches_fp32():
    from supereight2_amd import se

    ranges, poses = _street_scans()
    m32, s32 = _integrate(se, "f32", ranges, poses)
    m16, s16 = _integrate(se, "f16", ranges, poses)
    assert s16 == s32
    k32, ts32, t32, w32 = m32.dump_blocks()
    m32.close()
    k16, ts16, t16, w16 = m16.dump_blocks()
    m16.close()
    np.testing.assert_array_equal(k16, k32)
    np.testing.assert_array_equal(ts16, ts32)
    np.testing.assert_array_equal(w16, w32)
    # half an fp16 ulp of |tsdf| <= 1 per update, shrunk at each later update
    upd = w32 > 0
    assert np.abs(t16 - t32)[upd].max() < N_SCANS * 2.0**-12

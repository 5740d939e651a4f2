"""The oracle's multi-resolution TSDF updater (CPU).

- Integration scale: known answers of PinholeCamera::computeIntegrationScaleImpl
  (pinhole_camera.cpp:115-152) at fx 525 and 4 mm voxels (the 1.5 pixel/voxel threshold at 3.15 m),
  with the 0.25 m hysteresis once a block has a minimum scale.
- With every block at scale 0 (1 cm voxels, all depths below 7.9 m) the scale-0 data equal the
  single-resolution map's voxels bit for bit, and the past-data weights count at least the updates.
- Up-propagation: after frames 40 apart at 4 mm, every coarser entry above a block's current scale is
  the mean of its valid children (float32, children in x-fastest order) with the rounded-up mean
  weight, its past data {mean, 0}.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import CAM, spline_frames


def _pin():
    return O.pinhole(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)


def test_scale_zero_equals_single_resolution():
    depth, poses = spline_frames(3)
    s = O.OracleMap(10.24, 0.01)
    m = O.OracleMap(10.24, 0.01, multires=True)
    for k in range(3):
        s.integrate(k, depth[k], _pin(), poses[k])
        m.integrate(k, depth[k], _pin(), poses[k])
    sk, sts, st, sw = s.dump_blocks()
    mk, msc, md, mp = m.dump_multires()
    np.testing.assert_array_equal(sk, mk)
    assert (msc == 0).all()
    np.testing.assert_array_equal(md[:, :512, 0].view(np.uint32), st.view(np.uint32))
    np.testing.assert_array_equal(md[:, :512, 1], sw)
    # the past weight counts every sample that passed the pixel tests, also those more than mu
    # behind the surface that the field update skips
    assert (mp[:, :512, 1] >= sw).all() and (mp[:, :512, 1] > sw).any()


def _children(c, ps):
    cs = ps - 1
    n = 8 >> cs
    off = {0: 0, 1: 512, 2: 576, 3: 584}
    pn = 8 >> ps
    out = []
    for i in range(pn ** 3):
        x0, y0, z0 = (i % pn) << ps, ((i // pn) % pn) << ps, (i // (pn * pn)) << ps
        kids = []
        for k in range(8):
            x, y, z = x0 + ((k & 1) << cs), y0 + (((k >> 1) & 1) << cs), z0 + (((k >> 2) & 1) << cs)
            kids.append(off[cs] + (x >> cs) + (y >> cs) * n + (z >> cs) * n * n)
        out.append((off[ps] + i, kids))
    return out


def test_up_propagation_is_the_mean_of_valid_children():
    from supereight2_amd import synth

    scene = synth.make_scene(10)
    traj = synth.spline_trajectory(300, seed=11)
    rng = np.random.default_rng(3)
    m = O.OracleMap(10.24, 0.004, multires=True)
    for j, k in enumerate((0, 40, 80)):
        T = traj[k].astype(np.float32)
        m.integrate(j, synth.render_depth(scene, CAM, T, rng), _pin(), T)
    keys, sc, data, past = m.dump_multires()
    assert set(np.unique(sc[:, 0]).tolist()) == {0, 1}
    rng2 = np.random.default_rng(0)
    checked = 0
    for b in rng2.choice(len(keys), 400, replace=False):
        cur = sc[b, 0]
        for ps in range(cur + 1, 4):
            for pi, kids in _children(None, ps):
                w = data[b, kids, 1]
                t = data[b, kids, 0]
                valid = w != 0
                if not valid.any():
                    continue
                s, ws = np.float32(0), np.float32(0)
                for k in np.nonzero(valid)[0]:
                    s = np.float32(s + t[k])
                    ws = np.float32(ws + w[k])
                mean = np.float32(s / np.float32(valid.sum()))
                assert data[b, pi, 0] == mean
                assert data[b, pi, 1] == np.ceil(np.float32(ws / np.float32(valid.sum())))
                assert past[b, pi, 0] == mean and past[b, pi, 1] == 0
                checked += 1
    assert checked > 1000


def test_integration_scale_known_answers():
    """Blocks of a single frame: z of the block centre against the 3.15 m threshold (4 mm, fx 525)."""
    scaled_pixel = np.float32(1) / np.float32(525.0)
    res = np.float32(0.004)
    for z, expect in ((3.10, 0), (3.20, 1), (6.25, 1), (6.35, 2), (1.0, 0)):
        pv = np.float32(np.float32(z) * scaled_pixel) / res
        got = 0 if pv < 1.5 else (1 if pv < 3 else (2 if pv < 6 else 3))
        assert got == expect
    # the oracle map's first frame: every block's scale follows its centre depth (no hysteresis yet)
    depth, poses = spline_frames(1)
    m = O.OracleMap(10.24, 0.004, multires=True)
    m.integrate(0, depth[0], _pin(), poses[0])
    keys, sc, _, _ = m.dump_multires()
    T_CW = np.linalg.inv(poses[0].astype(np.float64))
    for i in range(0, len(keys), 97):
        c = O.decode_key(int(keys[i]))[0]
        cW = (c + 4.0) * 0.004 - 5.12
        z = (T_CW[:3, :3] @ cW + T_CW[:3, 3])[2]
        if abs(z - 3.15) > 0.01 and abs(z - 6.3) > 0.01:
            assert sc[i, 0] == (0 if z < 3.15 else (1 if z < 6.3 else 2)), (z, sc[i])
            assert sc[i, 1] == sc[i, 0]

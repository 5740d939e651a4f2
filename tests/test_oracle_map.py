"""The oracle's map, allocation, update, query and raycast functions against the reference's
unit-test KATs and against closed forms / brute force that a plausible mistake would fail."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def test_octree_size_kat():
    # test/map/map/map_unittest.cpp:210-236
    data = [(12.8, 0.1, 128), (25.6, 0.1, 256), (51.2, 0.1, 512), (10.24, 0.01, 1024), (20.48, 0.01, 2048),
            (10.24, 0.02, 512), (20.48, 0.02, 1024), (10.0, 0.05, 256), (100.0, 0.05, 2048)]
    for dim, res, size in data:
        m = O.OracleMap(dim, res)
        assert m.size == size


def _interp_map():
    m = O.OracleMap(32.0, 1.0)
    bs = 8
    coords = [(0, 0, 0), (bs, 0, 0), (0, bs, 0), (bs, bs, 0), (0, 0, bs), (bs, 0, bs), (0, bs, bs), (bs, bs, bs)]
    for i, bc in enumerate(coords):
        got = m.allocate_blocks([bc])
        np.testing.assert_array_equal(got[0], bc)
        assert m.set_block_data(bc, np.full(512, i, np.float32), np.ones(512, np.float32)) == 0
    return m


def test_map_interpolation_kat():
    # test/map/map/map_unittest.cpp:127-208
    m = _interp_map()
    queries = [(-12, -12, -12), (-8, -12, -12), (-12, -8, -12), (-8, -8, -12), (-12, -12, -8), (-8, -12, -8),
               (-12, -8, -8), (-8, -8, -8)]
    want = [0, 0.5, 1, 1.5, 2, 2.5, 3, 3.5]
    for q, w in zip(queries, want):
        assert m.field_interp(np.array(q, np.float32)) == w
    assert m.field_interp(np.array([2, 2, 2], np.float32)) is None


def test_allocation_kat_block_keys():
    # test/map/octree_allocator/octree_allocator_unittest.cpp:280-378
    m = O.OracleMap(256.0, 1.0)
    vox = np.array([[33, 104, 55], [113, 144, 155], [233, 44, 255]], np.int32)
    got = m.allocate_blocks(vox)
    want = sorted([tuple(int(x) for x in (v >> 3) << 3) for v in vox], key=lambda c: O.encode_code(c))
    np.testing.assert_array_equal(got, np.array(want))
    keys, masks, _ = m.dump_nodes()
    ks = set(int(k) for k in keys)
    for v in vox:
        for s in range(4, 9):
            assert O.encode_key((v >> s) << s, s)[0] in ks
    # nodes = root + the ancestors of the 3 blocks at scales 4..7 (distinct per block here)
    assert m.num_nodes == 1 + 3 * 4
    vox2 = np.concatenate([vox, [[0, 0, 0]]]).astype(np.int32)
    np.testing.assert_array_equal(m.allocate_blocks(vox2, only_new=True), [[0, 0, 0]])
    assert len(m.allocate_blocks(vox2, only_new=True)) == 0
    assert len(m.allocate_blocks(vox2)) == 4


def test_child_masks_match_children():
    m = O.OracleMap(256.0, 1.0)
    rng = np.random.default_rng(3)
    vox = rng.integers(0, 256, (200, 3)).astype(np.int32)
    m.allocate_blocks(vox)
    keys, masks, _ = m.dump_nodes()
    key_set = set(int(k) for k in keys)
    blocks = set(O.encode_key((v >> 3) << 3, 3)[0] for v in vox)
    for k, mk in zip(keys, masks):
        c, s = O.decode_key(int(k))
        for i in range(8):
            cc = c + (1 << (s - 1)) * np.array([(i & 1) > 0, (i & 2) > 0, (i & 4) > 0])
            ck = O.encode_key(cc, s - 1)[0]
            present = (ck in blocks) if s - 1 == 3 else (ck in key_set)
            assert bool((int(mk) >> i) & 1) == present


def _rot(axis, angle):
    a = np.float32(angle)
    c, s = np.float32(np.cos(a)), np.float32(np.sin(a))
    if axis == "y":
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], np.float32)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]], np.float32)


def test_aabb_of_wall_kat():
    # test/map/map/map_aabb_unittest.cpp:12-60: a uniform 4 m depth image, T_WC = T_BS
    pin = O.pinhole(100, 100, 100.0, 100.0, 100 / 2 - 0.5, 100 / 2 - 0.5, 0.1, 10.0)
    m = O.OracleMap(12.8, 0.1)
    T = np.eye(4, dtype=np.float32)
    T[:3, :3] = _rot("y", math.pi / 2) @ _rot("z", -math.pi / 2)
    m.integrate(0, np.full((100, 100), 4.0, np.float32), pin, T)
    valid, lo, hi = m.aabb()
    assert valid
    res, mu, t_mw = 0.1, 0.8, 6.4
    amin = lo * res - t_mw
    amax = np.nextafter(((hi + 1) * res - t_mw).astype(np.float32), np.float32(-1e9))
    assert amin[0] == pytest.approx(4.0 - mu, abs=1e-5)
    assert amax[0] == pytest.approx(4.0 + mu, abs=1e-5)
    assert amin[2] == pytest.approx(amin[1], abs=1e-5) and amax[2] == pytest.approx(amax[1], abs=1e-5)

    f32 = np.float32

    def dim_to_blocks(d, bd):  # float arithmetic, as in the reference test
        return int(math.copysign(float(f32(abs(f32(d))) / f32(bd) + f32(0.5)), float(d)))

    hfov = f32(2.0) * f32(math.atan(f32(100) / (f32(2.0) * f32(100.0))))
    half_wall = f32(4.0) * f32(math.tan(hfov / f32(2)))
    block_dim = f32(0.1) * f32(8)
    hb = dim_to_blocks(half_wall, block_dim)
    assert hb == 3
    assert dim_to_blocks(amin[1], block_dim) == -hb
    assert dim_to_blocks(amax[1], block_dim) == hb


def _wall(D=2.0, frames=1, precision="f32"):
    pin = O.pinhole()
    m = O.OracleMap(10.24, 0.01, precision=precision)
    depth = np.full((480, 640), D, np.float32)
    for k in range(frames):
        m.integrate(k, depth, pin, np.eye(4))
    return m, pin


def test_wall_tsdf_closed_form():
    """Fronto-parallel wall at D, identity pose: tsdf = min((D - z)/z * |p| / mu, 1) for every
    voxel that projects into the image with sdf >= -mu; weight = number of frames."""
    D, mu, res, t = 2.0, 0.08, 0.01, 5.12
    for frames in (1, 3):
        m, pin = _wall(D, frames)
        keys, ts, tsdf, w = m.dump_blocks()
        assert np.all(ts == frames - 1)
        idx = np.arange(512)
        off = np.stack([idx & 7, (idx >> 3) & 7, idx >> 6], -1)
        n_checked = 0
        for k, bt, bw in zip(keys, tsdf, w):
            c, _ = O.decode_key(int(k))
            p = (c[None, :] + off + 0.5) * res - t
            z = p[:, 2]
            nrm = np.linalg.norm(p, axis=1)
            u = 525.0 * p[:, 0] / z + 319.5
            v = 525.0 * p[:, 1] / z + 239.5
            inside = (z > 0) & (u >= -0.5) & (u < 639.5) & (v >= -0.5) & (v < 479.5) & (nrm <= 6.0 / (z / nrm))
            sdf = (D - z) / z * nrm
            upd = inside & (sdf >= -mu)
            # stay away from the projection / truncation boundaries where fp32 rounding decides
            safe = upd & (np.abs(sdf + mu) > 1e-4) & (np.abs(u + 0.5) > 1e-3) & (np.abs(u - 639.5) > 1e-3) \
                & (np.abs(v + 0.5) > 1e-3) & (np.abs(v - 479.5) > 1e-3)
            want = np.minimum(sdf / mu, 1.0)
            np.testing.assert_allclose(bt[safe], want[safe], atol=1.25e-4)  # |dtsdf| * mu <= 1e-5 m
            assert np.all(bw[safe] == frames)
            clear = ~inside | (sdf < -mu - 1e-4)
            assert np.all(bw[clear] == 0) and np.all(bt[clear] == 1.0)
            n_checked += int(safe.sum())
        assert n_checked > 100000


def test_wall_fp64_twin_matches_closed_form_tightly():
    # the twin receives the same float32 map parameters (res, dim) as the fp32 path
    D = 2.0
    res = float(np.float32(0.01))
    mu = res * 8.0
    t = float(np.float32(np.float32(10.24) / np.float32(2.0)))
    m, _ = _wall(D, 1, precision="f64")
    keys, _, tsdf, w = m.dump_blocks()
    idx = np.arange(512)
    off = np.stack([idx & 7, (idx >> 3) & 7, idx >> 6], -1)
    err = []
    for k, bt, bw in zip(keys[:400], tsdf[:400], w[:400]):
        c, _ = O.decode_key(int(k))
        p = (c[None, :] + off + 0.5) * res - t
        z = p[:, 2]
        nrm = np.linalg.norm(p, axis=1)
        sdf = (D - z) / z * nrm
        sel = bw == 1
        err.append(np.abs(bt[sel] - np.minimum(sdf[sel] / mu, 1.0)))
    assert np.concatenate(err).max() < 1e-12


def test_wall_raycast_hits_plane():
    D = 2.0
    m, pin = _wall(D, 2)
    pts, nrm, scale = m.raycast(pin, np.eye(4, dtype=np.float32))
    hit = np.any(pts != 0, axis=-1)
    assert hit[100:380, 100:540].all()
    # the linear zero crossing mixes raw nearest-voxel and interpolated samples
    # (raycaster_impl.hpp:497-526), so a few rays land up to a fraction of a voxel off; most are
    # exact to fp32 rounding
    err = np.abs(pts[hit][:, 2] - D)
    assert np.median(err) < 1e-5
    assert np.abs(pts[20:460, 20:620][..., 2] - D).max() < 2e-3
    assert err.max() < 5e-3
    good = np.any(nrm != 0, axis=-1)
    assert good[100:380, 100:540].mean() > 0.99
    np.testing.assert_allclose(nrm[good], np.broadcast_to([0, 0, -1], nrm[good].shape), atol=1e-3)
    assert np.all(scale[hit] == 0)


def _linear_field_map(a, b, precision="f32"):
    """Blocks 8..31^3 voxels filled with tsdf = a . v + b (v = voxel index), weight 1."""
    m = O.OracleMap(32.0, 1.0, precision=precision)
    dt = np.float64 if precision == "f64" else np.float32
    idx = np.arange(512)
    off = np.stack([idx & 7, (idx >> 3) & 7, idx >> 6], -1)
    for bx in (8, 16):
        for by in (8, 16):
            for bz in (8, 16):
                bc = np.array([bx, by, bz])
                m.allocate_blocks([bc])
                v = bc[None, :] + off
                m.set_block_data(bc, (v @ a + b).astype(dt), np.ones(512, dt))
    return m


def test_gradient_and_interp_of_linear_field():
    """Central differences and trilinear interpolation are exact on a linear field, so any wrong
    sample index, sign or blend weight shows up."""
    a = np.array([0.01, -0.02, 0.03])
    m = _linear_field_map(a, 0.1, precision="f64")
    rng = np.random.default_rng(1)
    for _ in range(200):
        vq = rng.uniform(10.6, 21.4, 3)  # keeps the 4^3 gradient stencil inside the blocks
        pW = vq - 16.0                   # T_MW = 16, res = 1
        g = m.field_grad(pW)
        assert g is not None
        np.testing.assert_allclose(g, a, atol=1e-12)
        f = m.field_interp(pW)
        assert f == pytest.approx((vq - 0.5) @ a + 0.1, abs=1e-12)
    # a stencil reaching an unallocated block gives no gradient
    assert m.field_grad(np.array([8.5 - 16.0, 16.0 - 16.0, 16.0 - 16.0])) is None


def test_first_block_matches_brute_force():
    """Entry t of the first allocated block along a ray vs a float64 ray/box search."""
    m = O.OracleMap(2.56, 0.01)  # 256^3 voxels
    rng = np.random.default_rng(7)
    vox = rng.integers(0, 256, (1500, 3)).astype(np.int32)
    blocks = m.allocate_blocks(vox)
    res, t_mw = 0.01, 1.28
    lo = blocks * res - t_mw
    hi = (blocks + 8) * res - t_mw
    n_checked = 0
    for _ in range(400):
        o = rng.uniform(-1.2, 1.2, 3).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        t0, t1 = O.OracleMap.first_block(m, o, d, 0.0, 10.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            ta = (lo - o.astype(np.float64)) / d.astype(np.float64)
            tb = (hi - o.astype(np.float64)) / d.astype(np.float64)
        tn = np.max(np.minimum(ta, tb), axis=1)
        tf = np.min(np.maximum(ta, tb), axis=1)
        ok = (tf > np.maximum(tn, 0.0))
        if not ok.any():
            continue
        entry = np.maximum(tn[ok], 0.0).min()
        # skip near-ties between two boxes, where either answer is a correct traversal
        cand = np.sort(np.maximum(tn[ok], 0.0))
        if len(cand) > 1 and cand[1] - cand[0] < 1e-4:
            continue
        assert t0 == pytest.approx(entry, abs=2e-5), (o, d)
        n_checked += 1
    assert n_checked > 100


def test_nan_depth_sample_propagates_as_in_reference():
    """A NaN depth sample passes the near-plane test (NaN < near is false, singleres_tsdf_updater_impl
    .hpp:71) and std::min(sdf / mu, 1) keeps the NaN quotient (data_field_impl.hpp:45), so every
    voxel that projects onto that pixel gets a NaN TSDF with its weight counted; an fmin-style
    minimum (which drops the NaN) would leave those voxels at tsdf 1 instead.  The carver skips the
    NaN ray (its points fall outside the map), so only blocks allocated by valid neighbours hold
    such voxels."""
    pin = O.pinhole()
    depth = np.full((pin.height, pin.width), 2.0, np.float32)
    depth[200:280, 280:360] = np.nan
    m = O.OracleMap(10.24, 0.01)
    m.integrate(0, depth, pin, np.eye(4))
    keys, ts, tsdf, w = m.dump_blocks()
    nan = np.isnan(tsdf)
    assert nan.sum() > 1000
    np.testing.assert_array_equal(w[nan], 1.0)
    # the NaN voxels all lie in front of the wall, inside the NaN window's frustum
    idx = np.arange(512)
    off = np.stack([idx & 7, (idx >> 3) & 7, idx >> 6], -1)
    for k, bn in zip(keys, nan):
        if not bn.any():
            continue
        c, _ = O.decode_key(int(k))
        p = (c[None, :] + off + 0.5) * 0.01 - 5.12
        u = 525.0 * p[:, 0] / p[:, 2] + 319.5
        v = 525.0 * p[:, 1] / p[:, 2] + 239.5
        ui, vi = (u[bn] + 0.5).astype(int), (v[bn] + 0.5).astype(int)
        assert ((ui >= 280) & (ui < 360) & (vi >= 200) & (vi < 280)).all()


def _popcount_rows(bits):
    return np.array([sum(bin(int(w)).count("1") for w in row) for row in bits])


def test_raycast_touched_set_is_union_of_its_rays():
    """The raycast's read set (the voxels behind the raycast roofline's algorithmic bytes) over a
    16 x 12 window of the image equals the union of the read sets of its 192 rays cast one by one
    (a 1 x 1 camera whose principal point puts its pixel on the same ray), and the counts are the
    popcount of that set and the number of blocks it touches."""
    from supereight2_amd import synth

    cam = synth.Camera()
    scene = synth.make_scene(0)
    poses = synth.circle_trajectory(50)[:2]
    rng = np.random.default_rng(1000)
    m = O.OracleMap(10.24, 0.01)
    pin = O.pinhole()
    for k, T in enumerate(poses):
        m.integrate(k, synth.render_depth(scene, cam, T, rng), pin, T.astype(np.float32))
    T = poses[1].astype(np.float32)
    x0, y0, w, h = 300, 200, 16, 12
    win = O.pinhole(w, h, cam.fx, cam.fy, cam.cx - x0, cam.cy - y0, cam.near_plane, cam.far_plane)
    _, _, _, V, B, bits = m.raycast_touched(win, T, bits=True)
    union = np.zeros_like(bits)
    for y in range(h):
        for x in range(w):
            one = O.pinhole(1, 1, cam.fx, cam.fy, cam.cx - (x0 + x), cam.cy - (y0 + y), cam.near_plane, cam.far_plane)
            _, _, _, v1, b1, bits1 = m.raycast_touched(one, T, bits=True)
            assert 0 < v1 <= 512 * b1
            union |= bits1
    np.testing.assert_array_equal(bits, union)
    per_block = _popcount_rows(bits)
    assert V == per_block.sum() and B == int((per_block > 0).sum())
    assert V > 192  # every ray reads more than its hit voxel


@pytest.mark.parametrize("tsdf", [1.0, 0.25])
def test_raycast_touched_one_block_kat(tsdf):
    """One observed block of constant tsdf on an axis-aligned ray: the march samples it every
    max(tsdf * mu, res) = 8 voxels (tsdf 1) or 2 voxels (tsdf 0.25, above the interpolation band),
    so it reads 1 or 4 voxels of the block, starting at the voxel of the block entry sample (when
    that sample rounds into the block; otherwise it takes the unobserved large step of 8 voxels and
    lands on the block's last voxel column)."""
    m = O.OracleMap(32.0, 1.0)  # 32^3 voxels of 1 m, T_MW = 16 m
    assert m.size == 32
    m.allocate_blocks([[8, 8, 8]])
    assert m.set_block_data([8, 8, 8], np.full(512, tsdf), np.ones(512)) == 0
    # camera at world (-20, -4.3, -4.7) looking along +x: camera z -> world x
    T = np.array([[0, 0, 1, -20.0], [1, 0, 0, -4.3], [0, 1, 0, -4.7], [0, 0, 0, 1]], np.float32)
    one = O.pinhole(1, 1, 1.0, 1.0, 0.0, 0.0, 0.4, 60.0)
    _, _, _, V, B, bits = m.raycast_touched(one, T, bits=True)
    t_min, _ = m.first_block(T[:3, 3], np.array([1, 0, 0], np.float32), 0.4, 60.0)
    assert abs(t_min - 12.0) < 1e-4  # the block's x face at world -8
    x_entry = np.float32(np.float32(-20.0) + np.float32(t_min)) + np.float32(16.0)
    in_block = x_entry >= 8.0
    step = 8 if tsdf == 1.0 else 2
    xs = list(range(8, 16, step)) if in_block else [15]
    want = np.zeros(512, bool)
    for x in xs:
        want[(x - 8) + 8 * (11 - 8) + 64 * (11 - 8)] = True
    got = np.array([(int(bits[0][i >> 6]) >> (i & 63)) & 1 for i in range(512)], bool)
    print("entry x", x_entry, "voxels read", np.nonzero(got)[0])
    assert B == 1 and V == len(xs)
    np.testing.assert_array_equal(got, want)


def test_raycast_stats_count_the_march():
    """The per-ray march statistics (used to analyse the raycast's latency chain) add up: a hit
    ray's samples are the march's samples, the in-octant counters are a split of their parent
    counters, and the one-block case of the read-set test takes exactly its 4 observed samples."""
    m = O.OracleMap(32.0, 1.0)
    m.allocate_blocks([[8, 8, 8]])
    assert m.set_block_data([8, 8, 8], np.full(512, 0.25), np.ones(512)) == 0
    T = np.array([[0, 0, 1, -20.0], [1, 0, 0, -4.3], [0, 1, 0, -4.7], [0, 0, 0, 1]], np.float32)
    one = O.pinhole(1, 1, 1.0, 1.0, 0.0, 0.0, 0.4, 60.0)
    st = m.raycast_stats(one, T)[0, 0]
    # 4 observed samples in the block, then large steps through unallocated space out of the map
    assert st[4] == 4 and st[3] == 0 and st[5] == 0
    assert st[0] > 0 and st[6] <= st[2]


def test_half_round_is_binary16_nearest_even():
    """The oracle's FP16 storage rounding against numpy's float32 -> float16 conversion (IEEE
    round to nearest, ties to even): special values, the subnormal range, ties, and random values
    over the TSDF range and beyond."""
    rng = np.random.default_rng(3)
    special = np.array([0.0, -0.0, 1.0, -1.0, 0.5, 65504.0, 65520.0, -65520.0, 6.1e-5, 5.96e-8, 2.98e-8, 2.99e-8,
                        1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, np.inf, -np.inf], np.float32)
    vals = np.concatenate([special, rng.uniform(-1, 1, 3000).astype(np.float32),
                           (rng.standard_normal(1000) * 1e-5).astype(np.float32),
                           (rng.standard_normal(500) * 3e4).astype(np.float32)])
    with np.errstate(over="ignore"):
        want = vals.astype(np.float16).astype(np.float32)
    np.testing.assert_array_equal(O.half_round(vals), want)
    assert np.isnan(O.half_round(np.array([np.nan], np.float32)))[0]


def test_fp16_storage_map_stays_within_rounding_of_fp32():
    """Two frames into FP32 and FP16 oracle maps: the same blocks and weights, stored values on the
    binary16 grid, and every tsdf within n_updates * 2^-12 of the FP32 map's (half an ulp of
    [0.5, 1) per rounded write, shrunk by each later average)."""
    from tests.helpers import CAM, spline_frames

    depth, poses = spline_frames(2)
    pin = O.pinhole(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)
    m32, m16 = O.OracleMap(10.24, 0.01), O.OracleMap(10.24, 0.01, storage="f16")
    for k in range(2):
        m32.integrate(k, depth[k], pin, poses[k])
        m16.integrate(k, depth[k], pin, poses[k])
    k32, _, t32, w32 = m32.dump_blocks()
    k16, _, t16, w16 = m16.dump_blocks()
    np.testing.assert_array_equal(k32, k16)
    np.testing.assert_array_equal(w32, w16)
    obs = w16 > 0
    np.testing.assert_array_equal(O.half_round(t16[obs][::97]), t16[obs][::97])
    assert np.all(np.abs(t16[obs] - t32[obs]) <= w16[obs] * 2.0 ** -12)
    assert np.abs(t16[obs] - t32[obs]).max() > 0

"""se8_mesh (Map::mesh, single-resolution marching cubes) on the GPU against the oracle.

Faces are compared as a set (the reference's face order follows its OpenMP schedule, the GPU's
the block pool): every face's nine floats bit-exact, NaN vertices (from NaN tsdf samples, kept by
checkVertex as in the reference) compared as NaN.  Cases: the reference meshing unit test setup;
blocks of special values (+-inf, NaN, +-3e38, signed zeros, denormals) and unknown voxels;
maps integrated from the seeded tracking frames with FP32 storage (oracle integrated alongside) and
FP16 storage (oracle loaded with the GPU map's blocks); a non-identity T_OW; corrupted depth (NaN
voxels); an empty map; a face buffer too small (SE8_ERR_CAPACITY with the count); saveMesh of the
GPU mesh read back.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import CAM, spline_frames

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _canon(f):
    """(F, 3, 3) float32 -> (F, 9) uint32 rows sorted lexicographically, every NaN as one pattern."""
    a = np.ascontiguousarray(f, dtype=np.float32).reshape(-1, 9).copy()
    a[np.isnan(a)] = np.float32("nan")
    u = a.view(np.uint32)
    return u[np.lexsort(u.T[::-1])]


def _assert_same_faces(g, o):
    assert g.shape == o.shape, (g.shape, o.shape)
    np.testing.assert_array_equal(_canon(g), _canon(o))


def _sensor(se):
    return se.PinholeCamera(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)


def _pin():
    return O.pinhole(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)


def _block_coords(keys):
    return np.array([O.decode_key(int(k))[0] for k in keys], np.int32).reshape(-1, 3)


def test_mesh_reference_unittest_setup():
    from supereight2_amd import se

    t = np.array([1.0 if (i & 7) < 4 else -1.0 for i in range(512)], np.float32)
    w = np.ones(512, np.float32)
    gmap = se.TSDFMap(0.16, 0.01)
    bcs = np.array([[8, y, z] for y in (0, 8) for z in (0, 8)], np.int32)
    gmap.allocate_blocks(bcs)
    for bc in bcs:
        gmap.write_block(bc, t, w)
    omap = O.OracleMap(0.16, 0.01)
    omap.load_blocks(bcs, np.tile(t, (4, 1)), np.tile(w, (4, 1)))
    g = gmap.mesh().cpu().numpy()
    assert g.shape == (450, 3, 3)
    _assert_same_faces(g, omap.mesh())
    gmap.close()


def test_mesh_special_values_match_oracle():
    """Blocks filled with special tsdf values (+-inf, NaN, +-3e38, +-0, +-1e-45, ordinary values)
    and zero weights, at the map's upper corner too: the GPU keeps every triangle of a cell's case,
    the oracle applies checkVertex literally, and the two meshes are equal."""
    from supereight2_amd import se

    rng = np.random.default_rng(11)
    special = np.array([np.inf, -np.inf, np.nan, 3e38, -3e38, 0.0, -0.0, 1e-45, -1e-45, 0.5, -0.5, 1.0, -1.0],
                       np.float32)
    gmap, omap = se.TSDFMap(0.32, 0.01), O.OracleMap(0.32, 0.01)  # 32 voxels: blocks 0, 8, 16, 24
    bcs = np.array([[x, y, z] for x in (8, 16, 24) for y in (16, 24) for z in (0, 24)], np.int32)
    gmap.allocate_blocks(bcs)
    omap.allocate_blocks(bcs)
    for bc in bcs:
        t = rng.choice(special, 512).astype(np.float32)
        plain = rng.random(512) < 0.5
        t[plain] = rng.uniform(-1, 1, int(plain.sum())).astype(np.float32)
        w = (rng.random(512) < 0.9).astype(np.float32)
        gmap.write_block(bc, t, w)
        assert omap.set_block_data(bc, t, w) == 0
    g = gmap.mesh().cpu().numpy()
    o = omap.mesh()
    print("special-value faces", len(g), "with NaN", int(np.isnan(g).any(axis=(1, 2)).sum()), flush=True)
    assert len(g) > 1000 and np.isnan(g).any()
    _assert_same_faces(g, o)
    gmap.close()


@pytest.mark.parametrize("corrupt", [False, True])
def test_mesh_of_integrated_frames_matches_oracle(corrupt):
    from supereight2_amd import se

    depth, poses = spline_frames(4)
    sensor, pin = _sensor(se), _pin()
    gmap, omap = se.TSDFMap(10.24, 0.01), O.OracleMap(10.24, 0.01)
    integ = se.MapIntegrator(gmap)
    rng = np.random.default_rng(5)
    for k in range(4):
        d = depth[k].copy()
        if corrupt:
            d[rng.random(d.shape) < 0.02] = np.nan
        print(f"frame {k} (corrupt={corrupt})", flush=True)
        integ.integrateDepth(k, torch.from_numpy(d).cuda(), sensor, poses[k])
        omap.integrate(k, d, pin, poses[k])
    torch.cuda.synchronize()
    g = gmap.mesh().cpu().numpy()
    o = omap.mesh()
    print("faces", len(g), "NaN faces", int(np.isnan(g).any(axis=(1, 2)).sum()), flush=True)
    assert len(g) > 100000
    if corrupt:
        assert np.isnan(g).any()
    _assert_same_faces(g, o)
    T = np.array([[0.0, -1.0, 0.0, 0.25], [1.0, 0.0, 0.0, -1.5], [0.0, 0.0, 1.0, 2.0], [0, 0, 0, 1]], np.float32)
    _assert_same_faces(gmap.mesh(T).cpu().numpy(), omap.mesh(T))
    gmap.close()


def test_mesh_fp16_storage_matches_oracle_on_the_same_voxels():
    from supereight2_amd import se

    depth, poses = spline_frames(3)
    sensor = _sensor(se)
    gmap = se.TSDFMap(10.24, 0.01, storage="f16")
    integ = se.MapIntegrator(gmap)
    for k in range(3):
        integ.integrateDepth(k, torch.from_numpy(depth[k]).cuda(), sensor, poses[k])
    torch.cuda.synchronize()
    keys, _, tsdf, w = gmap.dump_blocks()
    omap = O.OracleMap(10.24, 0.01)
    omap.load_blocks(_block_coords(keys), tsdf, w)
    g = gmap.mesh().cpu().numpy()
    print("fp16 faces", len(g), flush=True)
    _assert_same_faces(g, omap.mesh())
    gmap.close()


def test_mesh_empty_map_capacity_and_save(tmp_path):
    from supereight2_amd import capi, se

    empty = se.TSDFMap(10.24, 0.01)
    assert empty.mesh().shape == (0, 3, 3)
    empty.close()
    depth, poses = spline_frames(1)
    gmap = se.TSDFMap(10.24, 0.01)
    se.MapIntegrator(gmap).integrateDepth(0, torch.from_numpy(depth[0]).cuda(), _sensor(se), poses[0])
    ref = gmap.mesh()
    F = ref.shape[0]
    small = torch.zeros((F // 2, 3, 3), dtype=torch.float32, device="cuda")
    n = C.c_int64()
    rc = gmap.lib.se8_mesh(gmap.h, None, C.c_void_p(small.data_ptr()), F // 2, C.byref(n),
                           C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == capi.SE8_ERR_CAPACITY and n.value == F
    assert float(small.abs().sum()) == 0.0  # nothing written
    big = torch.zeros((F + 7, 3, 3), dtype=torch.float32, device="cuda")
    out = gmap.mesh(out=big)  # one call: fits in the caller's buffer
    assert out.shape == (F, 3, 3) and torch.equal(out, ref)
    again = gmap.mesh(out=small)  # too small: counted, then written to a new tensor
    assert torch.equal(again, ref)
    p = tmp_path / "mesh.ply"
    assert gmap.saveMesh(str(p)) == 0
    lines = p.read_text().split("\n")
    assert lines[3] == f"element vertex {3 * F}" and lines[7] == f"element face {F}"
    v = np.loadtxt(lines[13:13 + 3 * F])
    np.testing.assert_allclose(v, ref.cpu().numpy().reshape(-1, 3), rtol=5e-6, atol=1e-6)
    gmap.close()

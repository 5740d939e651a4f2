"""Debug views on the GPU against the oracle: Map::structure (six quads per octant) and the field
slices of Map::saveFieldSlices, after four integrated tracking frames (FP32 map, oracle integrated
alongside), with identity and rotated T_OW; the VTK slice files and the structure mesh files read
back.

Structure quads are compared as a set (the reference's order follows its octree iterator), all
twelve floats bit-exact, with their scales; slices bit-exact for the three axes through a point on
the surface and all 1 for a plane outside the octree.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import CAM, spline_frames

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _maps(n_frames=4):
    from supereight2_amd import se

    depth, poses = spline_frames(n_frames)
    sensor = se.PinholeCamera(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)
    pin = O.pinhole(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)
    gmap, omap = se.TSDFMap(10.24, 0.01), O.OracleMap(10.24, 0.01)
    integ = se.MapIntegrator(gmap)
    for k in range(n_frames):
        print(f"frame {k}", flush=True)
        integ.integrateDepth(k, torch.from_numpy(depth[k]).cuda(), sensor, poses[k])
        omap.integrate(k, depth[k], pin, poses[k])
    torch.cuda.synchronize()
    return gmap, omap, poses


def _sorted_quads(q, sc):
    rows = np.concatenate([np.ascontiguousarray(q, np.float32).reshape(-1, 12).view(np.uint32),
                           sc.astype(np.uint32).reshape(-1, 1)], axis=1)
    return rows[np.lexsort(rows.T[::-1])]


def test_structure_matches_oracle():
    gmap, omap, _ = _maps()
    T = np.array([[0.0, 0.0, 1.0, 0.5], [0.0, 1.0, 0.0, -1.0], [-1.0, 0.0, 0.0, 2.0], [0, 0, 0, 1]], np.float32)
    for only_leaves in (True, False):
        for T_OW in (None, T):
            q, sc = gmap.structure(T_OW, only_leaves)
            oq, osc = omap.structure(T_OW, only_leaves)
            print(f"only_leaves={only_leaves} quads {q.shape[0]}", flush=True)
            assert q.shape[0] == oq.shape[0] > 0
            np.testing.assert_array_equal(_sorted_quads(q.cpu().numpy(), sc.cpu().numpy()), _sorted_quads(oq, osc))
    assert gmap.structure()[0].shape[0] == 6 * gmap.info().num_blocks
    gmap.close()


def test_empty_map_structure_is_the_root():
    from supereight2_amd import se

    m = se.TSDFMap(10.24, 0.01)
    q, sc = m.structure()
    oq, osc = O.OracleMap(10.24, 0.01).structure()
    np.testing.assert_array_equal(_sorted_quads(q.cpu().numpy(), sc.cpu().numpy()), _sorted_quads(oq, osc))
    m.close()


def test_field_slices_match_oracle(tmp_path):
    gmap, omap, poses = _maps()
    size = gmap.octree_size()
    # the centre of an observed voxel next to the surface (|tsdf| < 0.05)
    keys, _, tsdf, w = omap.dump_blocks()
    b, i = np.argwhere((np.abs(tsdf) < 0.05) & (w > 0))[0]
    c = O.decode_key(int(keys[b]))[0] + np.array([i & 7, (i >> 3) & 7, i >> 6])
    p_W = (c + 0.5) * 0.01 - 5.12
    v = ((p_W + 5.12) / 0.01).astype(np.int32)
    for axis in range(3):
        g = gmap.field_slice(axis, int(v[axis])).cpu().numpy()
        o = omap.field_slice(axis, int(v[axis]))
        np.testing.assert_array_equal(g.view(np.uint32), o.view(np.uint32))
        assert (g != 1).sum() > 100
    assert (gmap.field_slice(1, size).cpu().numpy() == 1).all()
    names = [str(tmp_path / f"slice_{a}.vtk") for a in "xyz"]
    assert gmap.saveFieldSlices(names[0], "", names[2], p_W.astype(np.float32)) == 0
    assert not (tmp_path / "slice_y.vtk").exists()
    pv = (p_W.astype(np.float32) + np.float32(5.12)) / np.float32(0.01)
    vi = pv.astype(np.int32)
    for axis, name in ((0, names[0]), (2, names[2])):
        lines = open(name).read().split("\n")
        assert lines[:4] == ["# vtk DataFile Version 1.0", "vtk mesh generated from KFusion", "ASCII",
                             "DATASET RECTILINEAR_GRID"]
        dims = [size, size, size]
        dims[axis] = 1
        assert lines[4] == "DIMENSIONS " + " ".join(str(d) for d in dims)
        coords = lines[6 + 2 * axis].split()
        assert coords == [str(vi[axis])]
        assert lines[11] == f"POINT_DATA {size * size}" and lines[12] == "SCALARS scalars float 1"
        assert lines[13] == "LOOKUP_TABLE default"
        vals = np.array(lines[14:14 + size * size], np.float64)
        ref = omap.field_slice(axis, int(vi[axis])).reshape(-1)
        np.testing.assert_allclose(vals, ref, rtol=5e-6, atol=1e-7)
    gmap.close()


def test_save_structure_files(tmp_path):
    gmap, _, _ = _maps(1)
    q, sc = gmap.structure(only_leaves=False)
    Q = q.shape[0]
    p = tmp_path / "structure.ply"
    assert gmap.saveStructure(str(p), only_leaves=False) == 0
    lines = p.read_text().split("\n")
    assert lines[3] == f"element vertex {4 * Q}" and lines[7] == f"element face {Q}"
    faces = lines[13 + 4 * Q:13 + 5 * Q]
    n_grey = sum(f.endswith(" 179 179 179") for f in faces)
    assert n_grey == int((sc.cpu().numpy() == 7).sum()) > 0 and all(f.startswith("4 ") for f in faces)
    gmap.close()

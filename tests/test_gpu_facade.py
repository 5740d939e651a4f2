"""The header-only C++ facade (include/se8/supereight.hpp) driven like the reference app:
track -> integrateDepth -> raycast_volume per frame, on synthetic depth of a room corner, then
integrateDepth with an OusterLidar over a ground plane and a wall."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "tests", "cpp", "_build", "facade_app")


def test_facade_app_tracks_and_maps(tmp_path):
    assert os.path.exists(APP), "build with __graft_entry__.build() first"
    print("running facade_app", flush=True)
    res = subprocess.run([APP], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    print(res.stdout, res.stderr)
    assert res.returncode == 0
    frames = re.findall(r"frame (\d+) ok (\d) x (\S+) gt (\S+) hit_dist (\S+) depth_dist (\S+) scale (-?\d+)",
                        res.stdout)
    assert len(frames) == 6
    for f, ok, x, gt, hd, dd, sc in frames:
        assert ok == "1"
        assert abs(float(x) - float(gt)) < 2e-3
        assert abs(float(hd) - float(dd)) < 1e-2  # raycast hit at the measured surface
        assert sc == "0"
    # updated_octants (map_integrator.hpp:78-81): handles persist, so blocks in both frames share one
    m = re.search(r"updated frame4 (\d+) frame5 (\d+) shared (\d+) bad (\d+)", res.stdout)
    assert m and int(m.group(1)) > 1000 and int(m.group(2)) > 1000 and m.group(4) == "0"
    assert int(m.group(3)) > 0.5 * int(m.group(2))
    # TSDFMap<Res::Multi>: the first integration allocates and updates every block
    m = re.search(r"multires blocks (\d+) updated (\d+) quads (\d+)", res.stdout)
    assert m and int(m.group(1)) > 100 and m.group(1) == m.group(2) and int(m.group(3)) == 6 * int(m.group(1))
    m = re.search(r"interp (\d) (\S+) blocks (\d+)", res.stdout)
    assert m and m.group(1) == "1" and abs(float(m.group(2))) < 0.2 and int(m.group(3)) > 1000
    # Map::mesh / saveMesh: the room's surfaces, a .ply written in the working directory
    m = re.search(r"mesh faces (\d+) worst_plane_dist (\S+) saved (\d) bad_ext_threw (\d)", res.stdout)
    assert m and int(m.group(1)) > 10000 and float(m.group(2)) < 0.02 and m.group(3) == "0" and m.group(4) == "1"
    ply = (tmp_path / "facade_app_mesh.ply").read_text().split("\n")
    assert ply[0] == "ply" and ply[3] == f"element vertex {3 * int(m.group(1))}"
    # Map::structure (leaves: 6 quads per block), saveStructure and saveFieldSlices
    m = re.search(r"structure quads (\d+) leaves (\d+) saved (\d) slices (\d)", res.stdout)
    assert m and int(m.group(1)) == 6 * int(m.group(2)) and m.group(3) == "0" and m.group(4) == "0"
    assert (tmp_path / "facade_app_structure.vtk").exists() and (tmp_path / "facade_app_slice_z.vtk").exists()
    # lidar: the field at points on the ground plane and on the wall is the zero crossing, to within
    # the interpolation error of a 4 cm grid (mu = 16 cm, so 0.15 is 2.4 cm)
    m = re.search(r"lidar ground (\d) (\S+) wall (\d) (\S+) blocks (\d+)", res.stdout)
    assert m and m.group(1) == "1" and m.group(3) == "1" and int(m.group(5)) > 1000
    assert abs(float(m.group(2))) < 0.15 and abs(float(m.group(4))) < 0.15


def test_facade_app_reads_tum_sequence():
    """se::create_reader + se::preprocessor::downsample_depth + integrateDepth + mesh in C++ on the
    reference's 3-frame ICL-NUIM sequence (tests/golden/tum_icl3)."""
    seq = os.path.join(ROOT, "tests", "golden", "tum_icl3")
    print("running facade_app on the TUM sequence", flush=True)
    res = subprocess.run([APP, seq], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0
    m = re.search(r"tum frames (\d+) status (\d) valid (\S+) blocks (\d+) faces (\d+)", res.stdout)
    assert m and m.group(1) == "3" and m.group(2) == "3"  # the status after the last frame: error
    assert float(m.group(3)) > 0.95 and int(m.group(4)) > 5000 and int(m.group(5)) > 10000

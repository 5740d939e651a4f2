"""The allocation pipeline's kernels (supereight2_amd/csrc/se8_alloc.hpp) compiled unchanged as host
code and run on the CPU (tests/cpp/simt: one thread per lane, wave operations and workgroup barriers
emulated) under AddressSanitizer and UBSan, with the map's buffers allocated at exactly the sizes the
library gives them (plan_map_view).  Every out-of-bounds access of a buffer or of a workgroup's
__shared__ array is an error, and after every integration or explicit allocation the harness checks
the map's structure (keys, links, brick index, dense tables, AABB; tests/cpp/alloc_check.cpp).

* Claim path (maps of at most 4096 voxels a side: k_carve_claim + k_link_new) and sorted-key path
  (larger maps: k_alloc_rays, radix sort, k_build_level, k_build_blocks) over tracking frames of the
  seeded room: the block set and the update list equal the oracle's after every frame.
* The sorted-key build (build_level_one, k_rs_scatter, k_build_blocks) over the block lists of full
  640 x 480 oracle frames as explicit allocations (se8_allocate_blocks' path), unsorted and with
  repeats and blocks that exist, on both block tables: the block set equals the oracle's.
* Pool overflow on both paths (512 blocks): every access stays in bounds and the structure holds.
* Grids sized for 256 compute units (as on the device) instead of 4.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from supereight2_amd import synth
from tests.helpers import CAM

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "alloc_check.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "alloc_check")
DEPS = [SRC, os.path.join(ROOT, "tests", "cpp", "simt", "simt_host.hpp"),
        os.path.join(ROOT, "supereight2_amd", "csrc", "se8_alloc.hpp"),
        os.path.join(ROOT, "supereight2_amd", "csrc", "se8_internal.hpp"), os.path.join(ROOT, "include", "se8.h")]

SMALL = synth.Camera(width=160, height=120, fx=CAM.fx / 4, fy=CAM.fy / 4, cx=(CAM.cx + 0.5) / 4 - 0.5,
                     cy=(CAM.cy + 0.5) / 4 - 0.5)


@pytest.fixture(scope="module")
def alloc_check():
    if not os.path.exists(BIN) or os.path.getmtime(BIN) < max(os.path.getmtime(d) for d in DEPS):
        os.makedirs(os.path.dirname(BIN), exist_ok=True)
        cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
               "-fno-omit-frame-pointer", "-pthread", "-Wall", "-Wno-unknown-pragmas",
               "-I", os.path.join(ROOT, "tests", "cpp", "simt"), "-I", os.path.join(ROOT, "supereight2_amd", "csrc"),
               "-I", os.path.join(ROOT, "include"), SRC, "-o", BIN]
        res = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        assert res.returncode == 0, res.stderr
    return BIN


def _frames(cam, n, seed=0):
    scene = synth.make_scene(seed)
    poses = synth.spline_trajectory(300, seed=seed + 1)[:n]
    rng = np.random.default_rng(seed + 1000)
    return [(synth.render_depth(scene, cam, T, rng), np.asarray(T, np.float32)) for T in poses]


def _pin(cam):
    return O.pinhole(cam.width, cam.height, cam.fx, cam.fy, cam.cx, cam.cy, cam.near_plane, cam.far_plane)


def _write(path, dim, res, ops, max_blocks=0, cus=4):
    with open(path, "wb") as f:
        f.write(struct.pack("<5i", int(round(dim * 100)), int(round(res * 1e6)), max_blocks, cus, len(ops)))
        for op in ops:
            if op[0] == "integrate":
                _, cam, depth, T = op
                f.write(struct.pack("<3i", 0, cam.width, cam.height))
                f.write(struct.pack("<6f", cam.fx, cam.fy, cam.cx, cam.cy, cam.near_plane, cam.far_plane))
                f.write(np.ascontiguousarray(T, np.float32).tobytes())
                f.write(np.ascontiguousarray(depth, np.float32).tobytes())
            else:
                vox = np.ascontiguousarray(op[1], np.int32).reshape(-1, 3)
                f.write(struct.pack("<2i", 1, len(vox)))
                f.write(vox.tobytes())


def _run(exe, tmp_path, dim, res, ops, **kw):
    inp, out = tmp_path / "in.bin", tmp_path / "out.txt"
    _write(inp, dim, res, ops, **kw)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    res_ = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True, timeout=900, env=env)
    assert res_.returncode == 0 and res_.stdout.strip().endswith("OK"), (res_.stdout[-2000:], res_.stderr[-4000:])
    results, cur = [], None
    for line in out.read_text().split("\n"):
        if line.startswith("op "):
            f = line.split()
            cur = dict(blocks=int(f[3]), nodes=int(f[5]), overflow=int(f[7]), update=int(f[9]), b=[], u=[])
            results.append(cur)
        elif line.startswith("b "):
            cur["b"].append(int(line[2:]))
        elif line.startswith("u "):
            cur["u"].append(int(line[2:]))
    return results


def _check_frames_against_oracle(exe, tmp_path, dim, frames, cam):
    res = _run(exe, tmp_path, dim, 0.01, [("integrate", cam, d, T) for d, T in frames])
    omap = O.OracleMap(dim, 0.01)
    for k, ((d, T), r) in enumerate(zip(frames, res)):
        omap.integrate(k, d, _pin(cam), T)
        keys, _, _, _ = omap.dump_blocks(with_data=False)
        print(f"map {dim} m, {cam.width}x{cam.height} frame {k}: {r['blocks']} blocks, {r['nodes']} nodes, "
              f"{r['update']} updated")
        assert r["overflow"] == 0
        np.testing.assert_array_equal(np.array(r["b"], np.uint64), keys)
        np.testing.assert_array_equal(np.array(r["u"], np.uint64), omap.update_list())
        assert r["nodes"] == omap.num_nodes


def test_claim_path_frames(alloc_check, tmp_path):
    frames = _frames(SMALL, 4)
    _check_frames_against_oracle(alloc_check, tmp_path, 10.24, frames, SMALL)


def test_claim_path_full_frame(alloc_check, tmp_path):
    _check_frames_against_oracle(alloc_check, tmp_path, 10.24, _frames(CAM, 1), CAM)


def test_sorted_key_path_frames(alloc_check, tmp_path):
    frames = _frames(SMALL, 3)
    _check_frames_against_oracle(alloc_check, tmp_path, 81.92, frames, SMALL)


def _oracle_block_lists(n_frames):
    """Per full-size frame, voxel coordinates inside the blocks the oracle allocates for it: one
    random voxel per new block, every block of the previous list again, and a fifth of the list
    repeated, shuffled."""
    frames = _frames(CAM, n_frames)
    omap = O.OracleMap(10.24, 0.01)
    rng = np.random.default_rng(4)
    lists, prev = [], np.zeros((0, 3), np.int32)
    seen = set()
    for k, (d, T) in enumerate(frames):
        omap.integrate(k, d, _pin(CAM), T)
        keys, _, _, _ = omap.dump_blocks(with_data=False)
        new = [int(x) for x in keys if int(x) not in seen]
        seen.update(new)
        blocks = np.array([O.decode_key(x)[0] for x in new], np.int32).reshape(-1, 3)
        vox = blocks + rng.integers(0, 8, blocks.shape).astype(np.int32)
        lst = np.concatenate([vox, prev, vox[rng.choice(len(vox), len(vox) // 5)]])
        lists.append(lst[rng.permutation(len(lst))])
        prev = vox
    return lists


@pytest.mark.parametrize("dim", [10.24, 81.92])
def test_sorted_key_build_over_oracle_block_lists(alloc_check, tmp_path, dim):
    lists = _oracle_block_lists(3)
    res = _run(alloc_check, tmp_path, dim, 0.01, [("allocate", l) for l in lists])
    omap = O.OracleMap(dim, 0.01)
    for l, r in zip(lists, res):
        omap.allocate_blocks(l)
        keys, _, _, _ = omap.dump_blocks(with_data=False)
        print(f"map {dim} m: list of {len(l)} voxels -> {r['blocks']} blocks, {r['nodes']} nodes")
        assert r["overflow"] == 0
        np.testing.assert_array_equal(np.array(r["b"], np.uint64), keys)
        assert r["nodes"] == omap.num_nodes


@pytest.mark.parametrize("dim", [10.24, 81.92])
def test_pool_overflow_stays_in_bounds(alloc_check, tmp_path, dim):
    frames = _frames(SMALL, 2)
    lists = _oracle_block_lists(1)
    ops = [("integrate", SMALL, d, T) for d, T in frames] + [("allocate", lists[0])]
    res = _run(alloc_check, tmp_path, dim, 0.01, ops, max_blocks=512)
    print([(r["blocks"], r["nodes"], r["overflow"]) for r in res])
    assert res[0]["overflow"] == 1 and all(r["blocks"] == 512 for r in res)


@pytest.mark.parametrize("dim", [10.24, 81.92])
def test_grids_sized_for_the_device(alloc_check, tmp_path, dim):
    """256 compute units: k_frustum runs 1024 workgroups and k_link_new 256, most of them with no
    work (one frame, a few thousand blocks); the level build's grid covers the list."""
    frames = _frames(SMALL, 2)
    ops = [("integrate", SMALL, d, T) for d, T in frames]
    if dim > 10.24:
        ops = [("allocate", _oracle_block_lists(1)[0])]
    res = _run(alloc_check, tmp_path, dim, 0.01, ops, cus=256)
    assert all(r["overflow"] == 0 for r in res) and res[-1]["blocks"] > 1000

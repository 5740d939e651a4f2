"""Writes tests/golden/raycast_touched_spline300_seed10.json: per-frame algorithmic work of the
tracking workload, counted by the oracle (the CPU restatement of the reference) so that the
benchmark's byte counts do not depend on the GPU code they measure.

Workload: the benchmark's sequence, 300 frames of the seeded room (scene seed 10) along the smooth
spline (trajectory seed 11), 640x480, 1 cm voxels, 10.24 m map, truncation 8 voxels, ICP 10/5/4.
Depth is rendered by the numpy renderer with numpy noise (seed 1010); the benchmark renders the same
scene, trajectory and noise model on the GPU with torch's generator, so its frames are statistically
equivalent but not bit-identical.  The oracle runs the reference loop: frame 0 integrated at ground
truth, then track from the previous estimate against the previous raycast, integrate, raycast.

Per frame: the raycast's distinct voxel reads and the blocks holding them (raycast_volume's
getData samples, interpolation corners and gradient samples, raycaster_impl.hpp:456-627), the update
list length (frustum + new blocks), new blocks, carver candidates, valid pixels, and the Gauss-Newton
iterations run per pyramid level (finest first).

Run: python tests/golden/make_raycast_touched.py  (about 5 minutes on 8 cores)
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from supereight2_amd import synth  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "raycast_touched_spline300_seed10.json")
N_FRAMES = 300
SEED = 10


def main():
    O.set_threads(os.cpu_count() or 1)
    cam = synth.Camera()
    scene = synth.make_scene(SEED)
    gt = synth.spline_trajectory(N_FRAMES, seed=SEED + 1)
    rng = np.random.default_rng(SEED + 1000)
    pin = O.pinhole(cam.width, cam.height, cam.fx, cam.fy, cam.cx, cam.cy, cam.near_plane, cam.far_plane)
    m = O.OracleMap(10.24, 0.01)
    T = gt[0].astype(np.float32)
    pts = nrm = None
    frames = []
    t0 = time.time()
    for k in range(N_FRAMES):
        depth = synth.render_depth(scene, cam, gt[k], rng)
        rec = {"frame": k}
        if k > 0:
            ok, T, _, _, lev = O.track(depth, pin, T, pts, nrm, with_level_iterations=True)
            T = T.astype(np.float32)
            rec["track_ok"] = bool(ok)
            rec["icp_iterations"] = lev
        st = m.integrate(k, depth, pin, T)
        pts, nrm, _, V, B = m.raycast_touched(pin, T)
        rec.update(update_blocks=int(st["frustum"] + st["new"]), new_blocks=int(st["new"]),
                   candidates=int(st["candidates"]), valid_px=int(st["valid_px"]), raycast_voxels=V,
                   raycast_blocks=B, t_err_m=float(np.linalg.norm(T[:3, 3] - gt[k][:3, 3])))
        frames.append(rec)
        if k % 25 == 0:
            print(f"frame {k}: {rec} ({time.time() - t0:.0f} s)", flush=True)
    out = {
        "workload": "tracking sequence: room scene seed 10, spline trajectory seed 11, 300 frames, 640x480, "
                    "1 cm voxels, 10.24 m map, truncation 8 voxels, max weight 100, ICP 10/5/4",
        "frames_source": "numpy renderer (synth.render_depth), noise seed 1010",
        "counted_by": "oracle/se8_oracle.c (C restatement of the reference), fp32",
        "raycast_bytes": "24*W*H + s_v*raycast_voxels + 8*raycast_blocks (s_v = 8 FP32, 4 FP16)",
        "frames": frames,
    }
    with open(OUT, "w") as f:
        json.dump(out, f, indent=0, separators=(",", ":"))
    print("wrote", OUT)


if __name__ == "__main__":
    main()

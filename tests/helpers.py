"""Shared workload helpers for the parity tests (seeded synthetic frames, poses, cameras)."""
import functools

import numpy as np

from supereight2_amd import synth

CAM = synth.Camera()


@functools.lru_cache(maxsize=8)
def circle_frames(n_total=50, n=4, noise=True, seed=0):
    """First n frames of the seeded n_total-frame circle sequence (depth f32, poses f32)."""
    scene = synth.make_scene(seed)
    poses = synth.circle_trajectory(n_total)[:n]
    rng = np.random.default_rng(seed + 1000)
    depth = np.stack([synth.render_depth(scene, CAM, T, rng, noise=noise) for T in poses])
    return depth, np.stack(poses).astype(np.float32)


@functools.lru_cache(maxsize=4)
def spline_frames(n=6, seed=0, noise=True):
    scene = synth.make_scene(seed)
    poses = synth.spline_trajectory(300, seed=seed + 1)[:n]
    rng = np.random.default_rng(seed + 1000)
    depth = np.stack([synth.render_depth(scene, CAM, T, rng, noise=noise) for T in poses])
    return depth, np.stack(poses).astype(np.float32)

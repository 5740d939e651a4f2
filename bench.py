"""Benchmark: frames/s of track + integrate + raycast at 640x480, 1 cm voxels, on one MI355X.

Workload (one step = one frame of the KinectFusion loop of the reference app, main.cpp:98-162):
  Tracker::track against the previous frame's raycast (skipped for frame 0, which uses ground
  truth), MapIntegrator::integrateDepth at the tracked pose, raycast_volume at that pose.
  Sequence: 300 seeded synthetic frames of the room scene along the smooth spline trajectory
  (<= 2 cm, <= 1 degree per frame), 10.24 m map at 1 cm, truncation 8 voxels, max weight 100,
  pyramid iterations 10/5/4.  The first `warmup` frames are untimed; the next `steps` frames are
  timed.  Depth frames are resident in HBM before timing starts.

  Seed 10 (scene seed 10, trajectory seed 11) is the default because the reference algorithm
  tracks the whole sequence on it (the oracle: ATE 3.2 mm, no failed pose check).  With the
  Kinect noise model, sigma is about 2 cm at the 3.5 m typical of this room, so most per-pixel
  normals at 640x480 fail the 0.8 normal gate and the inlier fraction sits near the 0.15 check
  threshold; on seeds whose spheres leave long stretches of the path facing bare walls (0-5, 9,
  11) the reference tracker fails that check part-way and the pose freezes.

The FP32 path (reference arithmetic: fp32 TSDF + fp32 ICP reduction) gives `value`; the
FP16-TSDF + compensated-ICP variant runs on the same frames and is reported beside it.  The loop
keeps each frame's pose in device memory (track writes it, integrate and raycast read it), so the
host never waits inside the timed region; the same loop with host poses (the reference API's
Tracker::track returning the pose) is timed too and reported as config.host_pose_fps.

Prints one JSON line: the frames/s of the FP32 path as `value`, the FP16 variant beside it, the
roofline of the TSDF update kernel and the CPU-baseline rate of the oracle on a prefix of the frames.

`--workload street` runs the spinning-lidar mapping workload instead (one step = one
integrateDepth of a 1024x128 range scan; 200 scans along a 100 m street, 204.8 m map at 2 cm,
truncation 4 voxels) and prints its own line in the same format, in scans/s.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_FRAMES_DEFAULT = 300
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=290)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--seed", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-oracle error leg (about a minute of host time)")
    ap.add_argument("--cpu-frames", type=int, default=8, help="frames of the CPU baseline sample")
    ap.add_argument("--workload", choices=["room", "street"], default="room",
                    help="room: the 640x480 track+integrate+raycast loop (default); street: 200 scans of a "
                         "1024x128 spinning lidar along a 100 m street into a 204.8 m map at 2 cm (mapping only)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one 256 MiB device-to-device copy, the known-size reference that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters of a profiled run to bytes")
    return ap.parse_args()


COPY_CALIBRATION_BYTES = 256 << 20


def copy_calibration(torch, capi):
    """One plain device-to-device copy of COPY_CALIBRATION_BYTES through se8_copy."""
    lib = capi.load()
    src = torch.ones(COPY_CALIBRATION_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream().cuda_stream
    capi.check(lib.se8_copy(dst.data_ptr(), src.data_ptr(), COPY_CALIBRATION_BYTES, capi.COPY_D2D, s), "se8_copy")
    torch.cuda.synchronize()
    del src, dst


def ate(est, gt):
    e = np.array([np.linalg.norm(a[:3, 3] - b[:3, 3]) for a, b in zip(est, gt)])
    return float(np.sqrt(np.mean(e**2))) if len(e) else 0.0


def input_bench(depth_dev, torch, se):
    """preprocessor::downsample_depth fused with the RAW reader's conversion (uint16 millimetres ->
    metres) on frames of the workload, 640x480 -> 320x240 and 640x480 -> 640x480.  The launches are
    captured in a HIP graph (the host's per-call cost would otherwise dominate a kernel of a few
    microseconds) and the graph replayed between HIP events; the output of the replay is checked
    against an eager launch."""
    raw = torch.clamp(depth_dev[:8] * 1000.0, 0, 32767).round().to(torch.int16)
    out = {}
    for r in (2, 1):
        H, W = raw.shape[1] // r, raw.shape[2] // r
        dst = torch.empty((H, W), dtype=torch.float32, device="cuda")
        mp = torch.empty((H, W), dtype=torch.int64, device="cuda")
        n = 100
        ref = torch.empty_like(dst)
        se.downsample_depth_u16(raw[(n - 1) % 8], 1000.0, 1, ref)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        s_cap = torch.cuda.Stream()
        s_cap.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s_cap):
            se.downsample_depth_u16(raw[0], 1000.0, 1, dst, mp)  # warm-up outside the capture
            torch.cuda.synchronize()
            with torch.cuda.graph(g, stream=s_cap):
                for k in range(n):
                    se.downsample_depth_u16(raw[k % 8], 1000.0, 1, dst, mp)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            g.replay()
        e1.record()
        torch.cuda.synchronize()
        us = 1000.0 * e0.elapsed_time(e1) / (5 * n)
        nbytes = raw.shape[1] * raw.shape[2] * 2 + H * W * (4 + 8)
        out[f"ratio{r}"] = {"us": us, "bytes": nbytes, "gbs": nbytes / (us * 1e-6) / 1e9,
                            "replay_equal": bool(torch.equal(dst, ref))}
    return out


def mesh_bench(gmap, torch, se, cpu):
    """Map::mesh of the final map: wall time per call into a preallocated face buffer (count pass,
    scan, read-back of the total, emit pass), averaged over repeated calls; with cpu, the oracle's
    mesher (C restatement, OpenMP) on the same voxels."""
    ref = gmap.mesh()
    buf = torch.empty((ref.shape[0] + 1024, 3, 3), dtype=torch.float32, device="cuda")
    gmap.mesh(out=buf)
    torch.cuda.synchronize()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        gmap.mesh(out=buf)
    torch.cuda.synchronize()
    ms = 1000.0 * (time.perf_counter() - t0) / reps
    out = {"faces": int(ref.shape[0]), "ms_per_mesh": ms}
    # debug views of the same map: one field-slice plane (octree size squared lookups), kernel time
    # from a replayed HIP graph of 20 launches; Map::structure (leaves) wall time per call
    size = gmap.octree_size()
    sl = torch.empty((size, size), dtype=torch.float32, device="cuda")
    s_cap = torch.cuda.Stream()
    s_cap.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s_cap):
        gmap.lib.se8_field_slice(gmap.h, 2, size // 2, C.c_void_p(sl.data_ptr()), C.c_void_p(s_cap.cuda_stream))
        torch.cuda.synchronize()
        with torch.cuda.graph(g, stream=s_cap):
            for _ in range(20):
                gmap.lib.se8_field_slice(gmap.h, 2, size // 2, C.c_void_p(sl.data_ptr()),
                                         C.c_void_p(s_cap.cuda_stream))
    g.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        g.replay()
    e1.record()
    torch.cuda.synchronize()
    out["field_slice_us"] = 1000.0 * e0.elapsed_time(e1) / 100
    gmap.structure()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        q, _ = gmap.structure()
    torch.cuda.synchronize()
    out["structure_ms"] = 1000.0 * (time.perf_counter() - t0) / 10
    out["structure_quads"] = int(q.shape[0])
    if cpu:
        from oracle import oracle as O

        keys, _, tsdf, w = gmap.dump_blocks()
        coords = np.array([O.decode_key(int(k))[0] for k in keys], np.int32).reshape(-1, 3)
        omap = O.OracleMap(10.24, 0.01)
        omap.load_blocks(coords, tsdf, w)
        cores = os.cpu_count() or 1
        O.set_threads(cores)
        t0 = time.perf_counter()
        f = omap.mesh()
        out["cpu_ms_per_mesh"] = 1000.0 * (time.perf_counter() - t0)
        out["cpu_threads"] = cores
        out["cpu_faces_equal"] = int(len(f)) == out["faces"]
    return out


def run_pipeline(depth_dev, gt, n_frames, warmup, steps, storage, compensated, torch, se, sensor, timing="off",
                 device_pose=True, keep_points=False, dump=False, mesh=None, graph=False, prepare=False):
    """Runs the per-frame loop over all n_frames frames: frames [0, warmup) untimed, exactly `steps`
    frames [warmup, warmup + steps) timed, the rest of the sequence untimed after them (its time is
    kept separately for the whole-sequence rate).  With device_pose (the default), each frame's pose
    stays in device memory: track writes pose k from pose k - 1 and integrate / raycast read it there,
    so the host only enqueues work.  Without it, the loop is the reference API's (track returns the
    pose to the host, which passes it on).  `timing` is the library's stage-timing mode: "off" (no
    events, no host waits: the throughput figure), "stages" or "levels" (per-stage and per-pyramid-
    level device times from HIP events, accumulated over the timed frames).  keep_points keeps every
    frame's raycast vertex map on the device; dump returns the final map's blocks.  With graph, frames
    1.. are replays of one captured frame graph (se8_frame_graph_launch: the same kernels with the
    same arguments, dispatched back to back); it needs device poses, timing off and no kept points.
    With prepare (device poses), frame k's raycast also builds frame k+1's tracker pyramid in the same
    launch (se8_raycast_dev_prepare_track) and frame k+1's track skips its own, and each track's last
    launch also starts the frame's integration (se8_track_dev_prepare_integrate)."""
    H, W = sensor.height, sensor.width
    gmap = se.TSDFMap(10.24, 0.01, storage=storage)
    integ = se.MapIntegrator(gmap)
    tracker = se.Tracker(gmap, sensor, se.TrackerConfig(compensated=compensated))
    se.set_stage_timing(gmap, tracker, timing)
    if keep_points:
        pts = torch.zeros((n_frames, H, W, 3), dtype=torch.float32, device="cuda")
    else:
        pts = torch.zeros((2, H, W, 3), dtype=torch.float32, device="cuda")
    nrm = torch.zeros((2, H, W, 3), dtype=torch.float32, device="cuda")
    T = gt[0].astype(np.float32)
    T_dev = torch.zeros((n_frames, 16), dtype=torch.float32, device="cuda")
    T_dev[0] = torch.from_numpy(T.reshape(16))
    poses = []
    fg = None
    if graph:
        assert device_pose and timing == "off" and not keep_points
        # graph pair p is buffer 1 - p: launch k - 1 (frame k) raycasts into buffer k % 2 and tracks
        # against buffer (k + 1) % 2, as the per-call loop below does
        fg = se.FrameGraph(tracker, gmap, sensor, [pts[1], pts[0]], [nrm[1], nrm[0]])
    torch.cuda.synchronize()
    t0 = t1 = t2 = None
    for k in range(n_frames):
        if k == warmup:
            torch.cuda.synchronize()
            se.stage_times_reset(gmap, tracker)
            t0 = time.perf_counter()
        if k == warmup + steps:
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            st = se.stage_times(gmap, tracker)
        pk = k if keep_points else k % 2
        pp = (k - 1) if keep_points else (k + 1) % 2
        cur, prev = k % 2, (k + 1) % 2
        if fg is not None and k > 0:
            fg.launch(k, depth_dev[k], T_dev[k - 1], T_dev[k])
        elif device_pose:
            if k > 0 and prepare:
                se.track_device_prepare_integrate(tracker, depth_dev[k], T_dev[k - 1], T_dev[k], pts[pp], nrm[prev],
                                                  gmap)
            elif k > 0:
                tracker.track_device(depth_dev[k], T_dev[k - 1], T_dev[k], pts[pp], nrm[prev])
            integ.integrateDepth(k, depth_dev[k], sensor, T_dev[k])
            if prepare and k + 1 < n_frames:
                se.raycast_volume_prepare_track(gmap, sensor, T_dev[k], pts[pk], nrm[cur], tracker, depth_dev[k + 1])
            else:
                se.raycast_volume(gmap, sensor, T_dev[k], pts[pk], nrm[cur])
        else:
            if k > 0:
                _, T, _ = tracker.track(depth_dev[k], T, pts[pp], nrm[prev])
            integ.integrateDepth(k, depth_dev[k], sensor, T)
            se.raycast_volume(gmap, sensor, T, pts[pk], nrm[cur])
            poses.append(T.copy())
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    if t1 is None:
        t1 = t2
        st = se.stage_times(gmap, tracker)
    if device_pose:
        poses = list(T_dev.cpu().numpy().reshape(n_frames, 4, 4))
    info = gmap.info()
    mesh_out = mesh_bench(gmap, torch, se, mesh == "cpu") if mesh else None
    blocks = gmap.dump_blocks() if dump else None
    if fg is not None:
        fg.close()
    tracker.close()
    gmap.close()
    ni, nr, nt = max(st["n_integrations"], 1), max(st["n_raycasts"], 1), max(st["n_tracks"], 1)
    rest = n_frames - warmup - steps
    return {
        "fps": steps / (t1 - t0) if steps > 0 else 0.0,
        "seconds": t1 - t0,
        "sequence_fps": (n_frames - warmup) / (t2 - t0),
        "rest_frames": rest,
        "poses": poses,
        "points": pts if keep_points else None,
        "blocks_dump": blocks,
        "blocks": int(info.num_blocks),
        "stage_ms": {
            "fetch_frustum": st["sum_fetch_frustum"] / ni,
            "create_list": st["sum_create_list"] / ni,
            "allocate_list": st["sum_allocate_list"] / ni,
            "update": st["sum_update"] / ni,
            "raycast": st["sum_raycast"] / nr,
            "tracking": st["sum_tracking"] / nt,
        },
        "track_level_ms": [st["sum_track_level"][l] / max(st["n_track_level_launches"][l], 1) for l in range(3)],
        "counts": {"integrations": st["n_integrations"], "raycasts": st["n_raycasts"], "tracks": st["n_tracks"]},
        "update_blocks_per_frame": st["sum_update_blocks"] / ni,
        "candidates_per_frame": st["sum_candidates"] / ni,
        "mesh": mesh_out,
    }


MR_RES = 0.004
MR_STEP = 3


def run_multires(depth_dev, gt, n_frames, warmup, torch, se, sensor):
    """Multi-resolution TSDF integration (SE8_STORAGE_MULTIRES, 4 mm voxels, where the 1.5
    pixel/voxel threshold falls at 3.15 m inside the room) of every third frame of the workload at
    its ground-truth pose: frames/s and the mean per-frame stage times from the library's events."""
    gmap = se.TSDFMap(10.24, MR_RES, storage="multires", max_blocks=1 << 20)
    integ = se.MapIntegrator(gmap)
    idx = list(range(0, n_frames, MR_STEP))
    torch.cuda.synchronize()
    t0 = None
    for j, k in enumerate(idx):
        if j == warmup:
            torch.cuda.synchronize()
            se.stage_times_reset(gmap, None)
            t0 = time.perf_counter()
        integ.integrateDepth(j, depth_dev[k], sensor, gt[k].astype(np.float32))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = se.stage_times(gmap, None)
    ni = max(st["n_integrations"], 1)
    _, sc, _, _ = gmap.dump_blocks_multires()
    out = {"frames": len(idx) - warmup, "fps": (len(idx) - warmup) / dt, "blocks": int(len(sc)),
           "blocks_per_scale": np.bincount(sc[:, 0] + 1, minlength=5)[1:].tolist(),
           "stage_ms": {"fetch_frustum": st["sum_fetch_frustum"] / ni, "create_list": st["sum_create_list"] / ni,
                        "allocate_list": st["sum_allocate_list"] / ni, "update": st["sum_update"] / ni},
           "update_blocks_per_frame": st["sum_update_blocks"] / ni,
           "workload": f"every {MR_STEP}rd frame of the tracking sequence at its ground-truth pose, 10.24 m map at "
                       f"{MR_RES * 1000:.0f} mm, multi-resolution TSDF (scales 0-3)"}
    gmap.close()
    return out


def cpu_baseline_multires(depth_host, gt, sensor, n=4):
    """The oracle's multi-resolution updater (OpenMP) on frames 0, 3, .. of the same sequence."""
    from oracle import oracle as O

    O.set_threads(os.cpu_count() or 1)
    pin = O.pinhole(sensor.width, sensor.height, sensor.fx, sensor.fy, sensor.cx, sensor.cy, sensor.near_plane,
                    sensor.far_plane)
    m = O.OracleMap(10.24, MR_RES, multires=True)
    t0 = None
    for j in range(n):
        if j == 1:
            t0 = time.perf_counter()
        k = j * MR_STEP
        m.integrate(j, depth_host[k], pin, gt[k].astype(np.float32))
    return (n - 1) / (time.perf_counter() - t0)


def cpu_baseline(depth_host, gt, n_frames, sensor):
    """The oracle (C restatement of the reference, OpenMP) on a bounded prefix of the workload."""
    from oracle import oracle as O

    cores = os.cpu_count() or 1
    O.set_threads(cores)
    pin = O.pinhole(sensor.width, sensor.height, sensor.fx, sensor.fy, sensor.cx, sensor.cy, sensor.near_plane,
                    sensor.far_plane)
    m = O.OracleMap(10.24, 0.01)
    T = gt[0].astype(np.float32)
    pts = nrm = None
    # frame 0 (integrate + raycast) is warm-up; frames 1..n-1 are timed
    t0 = None
    for k in range(n_frames):
        if k == 1:
            t0 = time.perf_counter()
        if k > 0:
            _, T, _, _ = O.track(depth_host[k], pin, T, pts, nrm)
            T = T.astype(np.float32)
        m.integrate(k, depth_host[k], pin, T)
        pts, nrm, _ = m.raycast(pin, T)
    dt = time.perf_counter() - t0
    return (n_frames - 1) / dt, cores


STREET_SCANS = 200
STREET_MAX_BLOCKS = 6 << 20


def run_street(ranges_dev, poses, warmup, storage, torch, se, lidar):
    """Integrates every scan of the street sequence at its ground-truth pose (mapping only)."""
    gmap = se.TSDFMap(204.8, 0.02, truncation_boundary_factor=4.0, storage=storage, max_blocks=STREET_MAX_BLOCKS)
    integ = se.MapIntegrator(gmap)
    t0 = None
    for k in range(len(poses)):
        if k == warmup:
            torch.cuda.synchronize()
            se.stage_times_reset(gmap)
            t0 = time.perf_counter()
        integ.integrateDepth(k, ranges_dev[k], lidar, poses[k])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = se.stage_times(gmap)
    info = gmap.info()
    gmap.close()
    ni = max(st["n_integrations"], 1)
    timed = len(poses) - warmup
    return {
        "scans_per_s": timed / (t1 - t0) if timed > 0 else 0.0,
        "seconds": t1 - t0,
        "blocks": int(info.num_blocks),
        "nodes": int(info.num_nodes),
        "stage_ms": {
            "fetch_frustum": st["sum_fetch_frustum"] / ni,
            "create_list": st["sum_create_list"] / ni,
            "allocate_list": st["sum_allocate_list"] / ni,
            "update": st["sum_update"] / ni,
        },
        "update_blocks_per_scan": st["sum_update_blocks"] / ni,
        "candidates_per_scan": st["sum_candidates"] / ni,
        "integrations": st["n_integrations"],
    }


def street_main(args, torch, se, synth):
    """The lidar workload: one step = one integrateDepth of a 1024x128 range scan."""
    spec = synth.LidarSpec()
    street = synth.make_street(args.seed)
    n = min(args.warmup + args.steps, STREET_SCANS)
    warmup = min(args.warmup, n - 1)
    poses = synth.street_trajectory(STREET_SCANS)[:n].astype(np.float32)
    print(f"rendering {n} scans", file=sys.stderr, flush=True)
    ranges_dev = synth.render_street_ranges(street, spec, poses, seed=args.seed + 1, device="cuda")
    torch.cuda.synchronize()
    lidar = se.OusterLidar(spec.width, spec.height, spec.near_plane, spec.far_plane, spec.elevations_deg(),
                           spec.azimuths_deg())
    print("fp32 street", file=sys.stderr, flush=True)
    r32 = run_street(ranges_dev, poses, warmup, "f32", torch, se, lidar)
    r16 = None
    if not args.no_fp16:
        print("fp16 street", file=sys.stderr, flush=True)
        r16 = run_street(ranges_dev, poses, warmup, "f16", torch, se, lidar)
    lidar.close()

    def update_roofline(r, s_v):
        upd_ms = r["stage_ms"]["update"]
        nbytes = 2 * s_v * 512 * r["update_blocks_per_scan"] + 4 * spec.width * spec.height
        return nbytes, (nbytes / (upd_ms * 1e-3) / 1e9 if upd_ms > 0 else 0.0)

    b32, g32 = update_roofline(r32, 8)
    cpu = None
    if not args.no_cpu_baseline:
        print("cpu baseline", file=sys.stderr, flush=True)
        from oracle import oracle as O

        cores = os.cpu_count() or 1
        O.set_threads(cores)
        olid = O.OracleLidar(spec.width, spec.height, spec.near_plane, spec.far_plane, spec.elevations_deg(),
                             spec.azimuths_deg())
        om = O.OracleMap(204.8, 0.02, mu_factor=4.0)
        host = ranges_dev[: args.cpu_frames].cpu().numpy()
        t0 = None
        for k in range(args.cpu_frames):
            if k == 1:
                t0 = time.perf_counter()
            om.integrate_lidar(k, host[k], olid, poses[k])
        cpu = {"value": (args.cpu_frames - 1) / (time.perf_counter() - t0), "unit": "scans/s", "cores": cores,
               "kind": "port", "sample": f"scans 1..{args.cpu_frames - 1} of the same sequence (scan 0 untimed), "
                                         f"integrate, OpenMP on {cores} threads, {cpu_model()}"}
    fp16 = None
    if r16 is not None:
        b16, g16 = update_roofline(r16, 4)
        fp16 = {"scans_per_s": r16["scans_per_s"], "blocks": r16["blocks"], "stage_ms": r16["stage_ms"],
                "update_gbs": g16}
    out = {
        "metric": "scans/s integrate @1024x128 lidar, 2cm, 204.8m map; FP16-TSDF vs FP32",
        "value": r32["scans_per_s"],
        "unit": "scans/s",
        "n_gpus": 1,
        "steps": n - warmup,
        "warmup": warmup,
        "ms_per_step": 1000.0 * r32["seconds"] / max(n - warmup, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded street (ground, 40 box buildings, 200 cylinders), analytic range, N(0, 2 cm) noise",
        "config": {"workload": "fp32 TSDF mapping, 1024x128 beam lidar +-22.5 deg, 60 m range, 200 scans 0.5 m "
                               "apart along a 100 m street, 204.8 m map at 2 cm, mu 4 voxels",
                   "scans": n, "blocks": r32["blocks"], "nodes": r32["nodes"], "stage_ms": r32["stage_ms"],
                   "update_blocks_per_scan": r32["update_blocks_per_scan"],
                   "candidates_per_scan": r32["candidates_per_scan"]},
        "fp16_tsdf": fp16,
        "fp16_over_fp32": (r16["scans_per_s"] / r32["scans_per_s"]) if r16 is not None else None,
        "roofline": {"bound": "hbm", "kernel": "k_update", "achieved": g32, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": g32 / HBM_PEAK_GBS, "traffic": None, "bytes_per_launch": b32},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


FIXTURE = os.path.join(ROOT, "tests", "golden", "raycast_touched_spline300_seed10.json")


def _tsdf_compare(dump, dump64, mu):
    """Final maps: voxels observed in both, split into those with the same weight (the same number of
    updates, where the difference is arithmetic and storage precision) and those whose weights
    differ (a band-edge decision went the other way)."""
    keys, _, tg, wg = dump
    keys64, _, t64, w64 = dump64
    common, ig, io = np.intersect1d(keys, keys64, return_indices=True)
    tg_c, wg_c, to_c, wo_c = tg[ig], wg[ig], t64[io], w64[io]
    obs = (wg_c > 0) & (wo_c > 0)
    same = obs & (wg_c == wo_c)
    d_same = (tg_c[same].astype(np.float64) - to_c[same]) * mu
    d_all = (tg_c[obs].astype(np.float64) - to_c[obs]) * mu
    return {"tsdf_rms_m": float(np.sqrt(np.mean(d_same ** 2))) if d_same.size else None,
            "tsdf_max_m": float(np.abs(d_same).max()) if d_same.size else None,
            "tsdf_voxels": int(same.sum()),
            "tsdf_voxels_weight_differs": int((obs & (wg_c != wo_c)).sum()),
            "tsdf_rms_all_observed_m": float(np.sqrt(np.mean(d_all ** 2))) if d_all.size else None,
            "blocks_only_gpu": int(len(keys) - len(common)), "blocks_only_fp64": int(len(keys64) - len(common))}


class _VertexErr:
    def __init__(self):
        self.vmax, self.sq, self.cnt, self.p999, self.mismatch, self.frames = 0.0, 0.0, 0, [], 0, 0

    def add(self, g, p64):
        hg, h64 = np.any(g != 0, axis=-1), np.any(p64 != 0, axis=-1)
        both = hg & h64
        e = np.linalg.norm(g[both].astype(np.float64) - p64[both], axis=-1)
        if e.size:
            self.vmax = max(self.vmax, float(e.max()))
            self.p999.append(float(np.quantile(e, 0.999)))
            self.sq += float(np.sum(e * e))
            self.cnt += int(e.size)
        self.mismatch += int(np.sum(hg != h64))
        self.frames += 1

    def out(self):
        return {"vertex_max_m": self.vmax, "vertex_rms_m": float(np.sqrt(self.sq / max(self.cnt, 1))),
                "vertex_p999_max_m": float(max(self.p999)) if self.p999 else None,
                "hit_mismatch_px_per_frame": self.mismatch / max(self.frames, 1), "frames_compared": self.frames}


def fp64_errors(depth_host, gt, runs, sensor, mu, log_every=50, raycast_every=10):
    """Errors of each GPU run against the FP64 oracle (the C restatement in double precision) on the
    same frames, two ways:
    * same_poses: the FP64 oracle integrates every frame at the pose the GPU run used (its own
      tracked pose) and raycasts at it every `raycast_every` frames: the TSDF of the final maps and
      the vertex maps differ only by arithmetic and storage precision;
    * own_tracking: the FP64 oracle runs the whole reference loop itself (track from the previous
      estimate, integrate, raycast): the poses of the two runs, every frame's vertex map and the
      final TSDF, which then also carry the two trajectories' difference."""
    from oracle import oracle as O
    from supereight2_amd import synth

    O.set_threads(os.cpu_count() or 1)
    pin = O.pinhole(sensor.width, sensor.height, sensor.fx, sensor.fy, sensor.cx, sensor.cy, sensor.near_plane,
                    sensor.far_plane)
    n = len(depth_host)
    t0 = time.perf_counter()
    out = {"frames": n, "threads": os.cpu_count() or 1}
    res = {name: {} for name in runs}
    for name, r in runs.items():
        m = O.OracleMap(10.24, 0.01, precision="f64")
        ve = _VertexErr()
        for k in range(n):
            T = np.asarray(r["poses"][k], np.float32).astype(np.float64)
            m.integrate(k, depth_host[k], pin, T)
            if k % raycast_every == raycast_every - 1 or k == n - 1:
                p64, _, _ = m.raycast(pin, T)
                ve.add(r["points"][k].cpu().numpy(), p64)
            if log_every and k % log_every == 0:
                print(f"fp64 oracle at the {name} poses, frame {k} ({time.perf_counter() - t0:.0f} s)",
                      file=sys.stderr, flush=True)
        res[name]["same_poses"] = dict(_tsdf_compare(r["blocks_dump"], m.dump_blocks(), mu), **ve.out())
        del m
    m = O.OracleMap(10.24, 0.01, precision="f64")
    T = gt[0].astype(np.float64)
    p64 = n64 = None
    poses64 = []
    ves = {name: _VertexErr() for name in runs}
    for k in range(n):
        if k > 0:
            _, T, _, _ = O.track(depth_host[k], pin, T, p64, n64, precision="f64")
        m.integrate(k, depth_host[k], pin, T)
        p64, n64, _ = m.raycast(pin, T)
        poses64.append(T.copy())
        for name, r in runs.items():
            ves[name].add(r["points"][k].cpu().numpy(), p64)
        if log_every and k % log_every == 0:
            print(f"fp64 oracle tracking, frame {k} ({time.perf_counter() - t0:.0f} s)", file=sys.stderr, flush=True)
    dump64 = m.dump_blocks()
    out["ate64_m"] = ate(poses64, gt)
    for name, r in runs.items():
        pg = r["poses"]
        dt = [float(np.linalg.norm(np.asarray(a)[:3, 3] - b[:3, 3])) for a, b in zip(pg, poses64)]
        dr = [synth.rotation_angle_deg(np.asarray(a, np.float64)[:3, :3].T @ np.asarray(b, np.float64)[:3, :3])
              for a, b in zip(pg, poses64)]
        res[name]["own_tracking"] = dict(_tsdf_compare(r["blocks_dump"], dump64, mu), **ves[name].out(),
                                         pose_max_m=float(max(dt)), pose_max_deg=float(max(dr)),
                                         pose_final_m=dt[-1], ate_minus_fp64_m=ate(pg, gt) - out["ate64_m"])
        out[name] = res[name]
    out["seconds"] = time.perf_counter() - t0
    return out


def rooflines(r, fixture, cam, warmup, steps, s_v):
    """Algorithmic bytes (definitions in the workload fixture's header and DESIGN) over measured
    device time for the raycast, each ICP level, the update and the whole frame."""
    W, H = cam.width, cam.height
    fr = fixture["frames"][warmup: warmup + steps]
    ray_bytes = float(np.mean([24 * W * H + s_v * f["raycast_voxels"] + 8 * f["raycast_blocks"] for f in fr]))
    out = {}
    st = r["stage_ms"]

    def entry(kernel, nbytes, ms):
        gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        return {"bound": "hbm", "kernel": kernel, "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": gbs / HBM_PEAK_GBS, "bytes_per_launch": nbytes, "us_per_launch": ms * 1000.0}

    out["k_raycast"] = entry("k_raycast", ray_bytes, st["raycast"])
    upd_bytes = 2 * s_v * 512 * r["update_blocks_per_frame"] + 4 * W * H
    out["k_update"] = entry("k_update", upd_bytes, st["update"])
    icp_iters = np.mean([f["icp_iterations"] for f in fr if "icp_iterations" in f], axis=0)
    icp_frame_bytes = 0.0
    for l in range(3):
        n_l = (W >> l) * (H >> l)
        out[f"k_icp_iter_level{l}"] = entry(f"k_icp_iter (level {l})", 48.0 * n_l, r["track_level_ms"][l])
        icp_frame_bytes += icp_iters[l] * 48.0 * n_l
    cand = r["candidates_per_frame"]
    new_blocks = float(np.mean([f["new_blocks"] for f in fr]))
    alloc_bytes = 4 * W * H + 16 * cand + 16 * new_blocks
    frame_bytes = alloc_bytes + upd_bytes + ray_bytes + icp_frame_bytes + 4 * W * H
    out["frame"] = {"bound": "hbm", "achieved": frame_bytes * r["fps"] / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": frame_bytes * r["fps"] / 1e9 / HBM_PEAK_GBS, "bytes_per_frame": frame_bytes,
                    "parts": {"allocation": alloc_bytes, "update": upd_bytes, "raycast": ray_bytes,
                              "icp": icp_frame_bytes, "pyramid": 4 * W * H},
                    "icp_iterations_per_level": icp_iters.tolist()}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported: the path runs on a single GPU.",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from supereight2_amd import capi, se, synth

    capi.load()
    if args.copy_calibration:
        copy_calibration(torch, capi)
    if args.workload == "street":
        street_main(args, torch, se, synth)
        return
    cam = synth.Camera()
    sensor = se.PinholeCamera(cam.width, cam.height, cam.fx, cam.fy, cam.cx, cam.cy, cam.near_plane, cam.far_plane)
    # the whole 300-frame sequence always runs; exactly `steps` frames after `warmup` are timed
    n_frames = max(N_FRAMES_DEFAULT, args.warmup + args.steps)
    warmup, steps = args.warmup, args.steps
    scene = synth.make_scene(args.seed)
    gt = np.stack(synth.spline_trajectory(max(n_frames, 2), seed=args.seed + 1))[:n_frames]
    print(f"rendering {n_frames} frames", file=sys.stderr, flush=True)
    depth_dev = synth.render_sequence_torch(scene, cam, gt, seed=args.seed + 1000)
    torch.cuda.synchronize()
    with open(FIXTURE) as f:
        fixture = json.load(f)
    fp64 = not args.no_fp64 and args.seed == 10

    print("fp32 pipeline", file=sys.stderr, flush=True)
    r32 = run_pipeline(depth_dev, gt, n_frames, warmup, steps, "f32", False, torch, se, sensor, timing="off",
                       keep_points=fp64, dump=fp64, mesh="gpu" if args.no_cpu_baseline else "cpu", prepare=True)
    print("fp32 pipeline, pyramid in its own launch", file=sys.stderr, flush=True)
    r32u = run_pipeline(depth_dev, gt, n_frames, warmup, steps, "f32", False, torch, se, sensor, timing="off")
    unprepared_equal32 = bool(np.array_equal(np.stack(r32u["poses"]), np.stack(r32["poses"])))
    print("fp32 pipeline, frame graph", file=sys.stderr, flush=True)
    r32g = run_pipeline(depth_dev, gt, n_frames, warmup, steps, "f32", False, torch, se, sensor, timing="off",
                        graph=True)
    graph_equal32 = bool(np.array_equal(np.stack(r32g["poses"]), np.stack(r32["poses"])))
    ate32 = ate(r32["poses"], gt)
    print("fp32 pipeline, stage and level timing", file=sys.stderr, flush=True)
    r32t = run_pipeline(depth_dev, gt, n_frames, warmup, steps, "f32", False, torch, se, sensor, timing="levels")
    print("fp32 pipeline, host poses", file=sys.stderr, flush=True)
    r32h = run_pipeline(depth_dev, gt, n_frames, warmup, steps, "f32", False, torch, se, sensor, timing="off",
                        device_pose=False)
    r16 = r16g = r16t = r16p = None
    if not args.no_fp16:
        print("fp16 pipeline", file=sys.stderr, flush=True)
        r16 = run_pipeline(depth_dev, gt, n_frames, warmup, steps, "f16", True, torch, se, sensor, timing="off",
                           keep_points=fp64, dump=fp64, prepare=True)
        print("fp16 pipeline, frame graph", file=sys.stderr, flush=True)
        r16g = run_pipeline(depth_dev, gt, n_frames, warmup, steps, "f16", True, torch, se, sensor, timing="off",
                            graph=True)
        print("fp16 pipeline, stage and level timing", file=sys.stderr, flush=True)
        r16t = run_pipeline(depth_dev, gt, n_frames, warmup, steps, "f16", True, torch, se, sensor, timing="levels")
        print("fp16 storage, plain ICP", file=sys.stderr, flush=True)
        r16p = run_pipeline(depth_dev, gt, n_frames, warmup, steps, "f16", False, torch, se, sensor, timing="off")

    # stage times and level times come from the timed-event runs on the same frames
    r32["stage_ms"], r32["track_level_ms"] = r32t["stage_ms"], r32t["track_level_ms"]
    r32["update_blocks_per_frame"], r32["candidates_per_frame"] = r32t["update_blocks_per_frame"], \
        r32t["candidates_per_frame"]
    roof32 = rooflines(r32, fixture, cam, warmup, steps, 8)
    # HBM traffic per launch of the raycast and update kernels on this workload, measured by the
    # FETCH_SIZE and WRITE_SIZE passes recorded under profiles/ (scaled by a 256 MiB copy); null
    # when that record is absent
    for kern, fname in (("k_raycast", "k_raycast_traffic.json"), ("k_update", "k_update_traffic.json")):
        tpath = os.path.join(ROOT, "profiles", fname)
        roof32[kern]["traffic"] = None
        if os.path.exists(tpath):
            with open(tpath) as f:
                tr = json.load(f)
            roof32[kern]["traffic"] = tr["traffic_bytes_per_launch"]
            roof32[kern]["traffic_source"] = f"profiles/{fname}"

    print("multi-resolution TSDF", file=sys.stderr, flush=True)
    rmr = run_multires(depth_dev, gt, n_frames, 5, torch, se, sensor)
    if not args.no_cpu_baseline:
        rmr["cpu_fps"] = cpu_baseline_multires(depth_dev[: 3 * MR_STEP + 1].cpu().numpy(), gt, sensor)
        rmr["cpu_threads"] = os.cpu_count() or 1

    errors = None
    if fp64:
        print("fp64 oracle over the sequence", file=sys.stderr, flush=True)
        runs = {"fp32": r32}
        if r16 is not None:
            runs["fp16_tsdf"] = r16
        errors = fp64_errors(depth_dev.cpu().numpy(), gt, runs, sensor, mu=0.08)

    cpu = None
    if not args.no_cpu_baseline:
        print("cpu baseline", file=sys.stderr, flush=True)
        depth_host = depth_dev[: args.cpu_frames].cpu().numpy()
        cfps, cores = cpu_baseline(depth_host, gt, args.cpu_frames, sensor)
        cpu = {"value": cfps, "unit": "frames/s", "cores": cores, "kind": "port",
               "sample": f"frames 1..{args.cpu_frames - 1} of the same sequence (frame 0 untimed), "
                         f"track+integrate+raycast, OpenMP on {cores} threads, {cpu_model()}"}

    fp16 = None
    if r16 is not None:
        r16["stage_ms"], r16["track_level_ms"] = r16t["stage_ms"], r16t["track_level_ms"]
        r16["update_blocks_per_frame"], r16["candidates_per_frame"] = r16t["update_blocks_per_frame"], \
            r16t["candidates_per_frame"]
        roof16 = rooflines(r16, fixture, cam, warmup, steps, 4)
        fp16 = {"fps": r16["fps"], "sequence_fps": r16["sequence_fps"], "frame_graph_fps": r16g["fps"],
                "frame_graph_poses_equal": bool(np.array_equal(np.stack(r16g["poses"]), np.stack(r16["poses"]))),
                "ate_m": ate(r16["poses"], gt),
                "ate_minus_fp32_m": ate(r16["poses"], gt) - ate32, "blocks": r16["blocks"],
                "stage_ms": r16["stage_ms"], "track_level_us": [1000 * x for x in r16["track_level_ms"]],
                "stage_delta_ms_vs_fp32": {k: r16["stage_ms"][k] - r32["stage_ms"][k] for k in r32["stage_ms"]},
                "rooflines": {k: {"achieved": v["achieved"], "frac": v["frac"]} for k, v in roof16.items()},
                "storage_only": {"fps": r16p["fps"], "ate_m": ate(r16p["poses"], gt),
                                 "note": "FP16 TSDF storage with the plain fp32 ICP reduction"}}
        if errors is not None:
            fp16["fp64"] = errors["fp16_tsdf"]
    main_roof = dict(roof32["k_raycast"])
    out = {
        "metric": "frames/s integrate+raycast+ICP @640x480,1cm; FP16-TSDF vs FP32 at stated RMSE",
        "value": r32["fps"],
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": 1000.0 * r32["seconds"] / max(steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded room + 3 spheres, analytic depth with Kinect-like noise, 1% dropout",
        "config": {"workload": f"fp32 TSDF + ICP tracking on frames {warmup}..{warmup + steps - 1} of the "
                               f"{n_frames}-frame spline sequence (the rest runs untimed), 640x480, 1 cm voxels, "
                               f"10.24 m map, mu 8 voxels, max weight 100, ICP 10/5/4",
                   "frames": n_frames, "timed_frames": [warmup, warmup + steps - 1],
                   "sequence_fps": r32["sequence_fps"], "ate_m": ate32, "blocks": r32["blocks"],
                   "stage_ms": r32["stage_ms"], "track_level_us": [1000 * x for x in r32["track_level_ms"]],
                   "stage_note": "stage and level times from HIP events in a second run of the same frames; the "
                                 "timed run records no events",
                   "poses": "device-resident (track -> integrate -> raycast without host round trips)",
                   "frame_graph_fps": r32g["fps"], "frame_graph_poses_equal": graph_equal32,
                   "unprepared_fps": r32u["fps"], "unprepared_poses_equal": unprepared_equal32,
                   "prepare_note": "value: each frame's raycast also builds the next frame's tracker pyramid in the "
                                   "same launch (se8_raycast_dev_prepare_track) and each track's last launch starts "
                                   "the integration (se8_track_dev_prepare_integrate); unprepared_fps: the same loop "
                                   "with the pyramid and the frame start as launches of their own",
                   "frame_graph_note": "the same loop with frames 1.. replayed from one captured frame graph "
                                       "(se8_frame_graph_launch), the pyramid in the track's own launch; "
                                       "value is the per-call loop",
                   "host_pose_fps": r32h["fps"], "host_pose_ate_m": ate(r32h["poses"], gt),
                   "update_blocks_per_frame": r32["update_blocks_per_frame"],
                   "candidates_per_frame": r32["candidates_per_frame"]},
        "fp16_tsdf": fp16,
        "fp16_over_fp32": (r16["fps"] / r32["fps"]) if r16 is not None and r32["fps"] > 0 else None,
        "fp64": ({"fp32": errors["fp32"], "ate64_m": errors["ate64_m"], "frames": errors["frames"],
                  "seconds": errors["seconds"], "threads": errors["threads"],
                  "note": "errors against the FP64 oracle on the same frames. same_poses: FP64 integration at "
                          "the GPU run's own poses (final TSDF over voxels with the same weight in both maps, in "
                          "metres; vertex maps of every 10th frame at those poses, pixels both hit): arithmetic "
                          "and storage precision only. own_tracking: the FP64 oracle's own tracked run (poses, "
                          "every frame's vertex map, final TSDF), which adds the two trajectories' difference. "
                          "The FP16 path's errors are in fp16_tsdf.fp64"}
                 if errors is not None else None),
        "roofline": main_roof,
        "rooflines": roof32,
        "cpu_baseline": cpu,
        "mesh": dict(r32["mesh"], blocks=r32["blocks"],
                     note="Map::mesh (single-res marching cubes) of the final FP32 map, after the timed loop"),
        "multires": rmr,
        "downsample": dict(input_bench(depth_dev, torch, se),
                           note="downsample_depth fused with the uint16 mm -> m conversion, index map written; "
                                "kernel time from a replayed HIP graph of 100 launches"),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

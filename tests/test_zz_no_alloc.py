"""se8_icp_cost_grad_batched_noalloc (device inputs and outputs, partials in the library's fixed
scratch array, no allocation) against se8_icp_cost_grad_batched on the same inputs: cost, gradient,
J^T J and inlier counts bit for bit, on two calls in a row, the second with fewer poses; sizes
beyond the scratch array are refused before any GPU work."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import CAM, circle_frames

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _candidates(T, n, seed):
    from supereight2_amd import diff

    rng = np.random.default_rng(seed)
    tw = np.concatenate([rng.normal(0, 0.02, (n, 3)), rng.normal(0, np.radians(1.5), (n, 3))], 1)
    return np.stack([diff.se3_exp(x) @ T.astype(np.float64) for x in tw]).astype(np.float32)


def test_noalloc_matches_allocating_entry_point():
    from supereight2_amd import capi, diff, se

    lib = capi.load()
    depth, poses = circle_frames(50, 4)
    sensor = se.PinholeCamera(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)
    gmap = se.TSDFMap(10.24, 0.01)
    integ = se.MapIntegrator(gmap)
    for k in range(3):
        integ.integrateDepth(k, torch.from_numpy(depth[k]).cuda(), sensor, poses[k])
    ref_p = torch.zeros((CAM.height, CAM.width, 3), dtype=torch.float32, device="cuda")
    ref_n = torch.zeros_like(ref_p)
    se.raycast_volume(gmap, sensor, poses[2], ref_p, ref_n)
    tracker = se.Tracker(gmap, sensor)
    d = torch.from_numpy(depth[3]).cuda()
    stream = torch.cuda.current_stream()
    for call, n in enumerate((64, 17)):
        cands = _candidates(poses[3], n, 100 + call)
        print(f"call {call}: {n} poses", flush=True)
        cost, grad, jtj, inl = diff.icp_cost_grad(tracker, d, ref_p, ref_n, poses[2], cands)
        assert np.all(inl > 1000)
        p_dev = torch.from_numpy(cands.reshape(n, 16)).cuda()
        c_dev = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
        g_dev = torch.full((n, 6), -1.0, dtype=torch.float32, device="cuda")
        j_dev = torch.full((n, 21), -1.0, dtype=torch.float32, device="cuda")
        i_dev = torch.full((n,), -1, dtype=torch.int32, device="cuda")
        capi.check(lib.se8_icp_cost_grad_batched_noalloc(
            tracker.h, C.c_void_p(d.data_ptr()), C.c_void_p(ref_p.data_ptr()), C.c_void_p(ref_n.data_ptr()),
            capi.pose16(poses[2]), n, C.c_void_p(p_dev.data_ptr()), C.c_void_p(c_dev.data_ptr()),
            C.c_void_p(g_dev.data_ptr()), C.c_void_p(j_dev.data_ptr()), C.c_void_p(i_dev.data_ptr()),
            C.c_void_p(stream.cuda_stream)), "se8_icp_cost_grad_batched_noalloc")
        torch.cuda.synchronize()
        assert np.array_equal(c_dev.cpu().numpy(), cost)
        assert np.array_equal(g_dev.cpu().numpy(), grad)
        assert np.array_equal(j_dev.cpu().numpy(), jtj)
        assert np.array_equal(i_dev.cpu().numpy(), inl)
    # more poses than the scratch array holds: refused on the host, nothing launched
    big = torch.zeros((65, 16), dtype=torch.float32, device="cuda")
    out = torch.zeros((65, 6), dtype=torch.float32, device="cuda")
    rc = lib.se8_icp_cost_grad_batched_noalloc(
        tracker.h, C.c_void_p(d.data_ptr()), C.c_void_p(ref_p.data_ptr()), C.c_void_p(ref_n.data_ptr()),
        capi.pose16(poses[2]), 65, C.c_void_p(big.data_ptr()), C.c_void_p(out.data_ptr()), C.c_void_p(out.data_ptr()),
        None, None, C.c_void_p(stream.cuda_stream))
    assert rc == capi.SE8_ERR_INVALID and b"n_poses" in lib.se8_last_error()
    tracker.close()
    gmap.close()

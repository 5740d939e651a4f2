"""torch.autograd wrappers of the adjoint kernels (check harness only).

The twist parameterisation is local: for candidate poses T0_k and perturbations delta_k,
T_k = exp(delta_k) T0_k with delta = [v; w] (the reference's left update, icp.cpp:215-216); the
kernels return exact derivatives at delta = 0, so gradients are taken at delta = 0 and the poses
are updated outside autograd (as a Gauss-Newton or gradient step would)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .se import PinholeCamera, TSDFMap, Tracker, _dev_f32, _stream_handle


def hat(w):
    return np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])


def se3_exp(xi) -> np.ndarray:
    """Closed-form SE(3) exponential (float64) of xi = [v; w]."""
    xi = np.asarray(xi, np.float64)
    v, w = xi[:3], xi[3:]
    th = np.linalg.norm(w)
    W = hat(w)
    if th < 1e-12:
        R = np.eye(3) + W
        V = np.eye(3) + 0.5 * W
    else:
        A = np.sin(th) / th
        B = (1 - np.cos(th)) / th**2
        Cc = (th - np.sin(th)) / th**3
        R = np.eye(3) + A * W + B * W @ W
        V = np.eye(3) + B * W + Cc * W @ W
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = V @ v
    return T


def icp_cost_grad(tracker: Tracker, depth, ref_points, ref_normals, T_ref, poses, stream=None):
    """cost (n,), grad (n, 6), jtj (n, 21), inliers (n,) for poses (n, 4, 4)."""
    poses = np.ascontiguousarray(np.asarray(poses, np.float32).reshape(-1, 16))
    n = len(poses)
    cost = np.zeros(n, np.float32)
    grad = np.zeros((n, 6), np.float32)
    jtj = np.zeros((n, 21), np.float32)
    inl = np.zeros(n, np.int32)
    H, W = tracker.sensor.height, tracker.sensor.width
    capi.check(tracker.lib.se8_icp_cost_grad_batched(
        tracker.h, _dev_f32(depth, (H, W)), _dev_f32(ref_points, (H, W, 3)), _dev_f32(ref_normals, (H, W, 3)),
        capi.pose16(T_ref), n, capi.fptr(poses), capi.fptr(cost), capi.fptr(grad), capi.fptr(jtj), capi.iptr(inl),
        _stream_handle(stream)), "se8_icp_cost_grad_batched")
    return cost, grad, jtj, inl


def raycast_depth_grad(map_: TSDFMap, sensor: PinholeCamera, poses, stream=None):
    """depth (n, H, W), grad (n, H, W, 6), valid (n, H, W) as device tensors."""
    import torch

    poses = np.ascontiguousarray(np.asarray(poses, np.float32).reshape(-1, 16))
    n = len(poses)
    H, W = sensor.height, sensor.width
    depth = torch.zeros((n, H, W), dtype=torch.float32, device="cuda")
    grad = torch.zeros((n, H, W, 6), dtype=torch.float32, device="cuda")
    valid = torch.zeros((n, H, W), dtype=torch.uint8, device="cuda")
    capi.check(map_.lib.se8_raycast_depth_grad_batched(
        map_.h, C.byref(sensor.c()), n, capi.fptr(poses), C.c_void_p(depth.data_ptr()), C.c_void_p(grad.data_ptr()),
        C.c_void_p(valid.data_ptr()), _stream_handle(stream)), "se8_raycast_depth_grad_batched")
    return depth, grad, valid.bool()


def _autograd():
    import torch

    class PointToPlaneCost(torch.autograd.Function):
        """cost_k(delta_k) = 0.5 sum e_i^2 at exp(delta_k) T0_k (exact gradient at delta = 0)."""

        @staticmethod
        def forward(ctx, delta, tracker, depth, ref_points, ref_normals, T_ref, T0):
            d = delta.detach().cpu().double().numpy()
            poses = np.stack([se3_exp(d[k]) @ T0[k] for k in range(len(T0))])
            cost, grad, _, _ = icp_cost_grad(tracker, depth, ref_points, ref_normals, T_ref, poses)
            ctx.save_for_backward(torch.from_numpy(grad).to(delta.device, delta.dtype))
            return torch.from_numpy(cost).to(delta.device, delta.dtype)

        @staticmethod
        def backward(ctx, g):
            (grad,) = ctx.saved_tensors
            return g[:, None] * grad, None, None, None, None, None, None

    class RaycastDepth(torch.autograd.Function):
        """depth_k(delta_k) of the refined raycast at exp(delta_k) T0_k (exact gradient at delta = 0)."""

        @staticmethod
        def forward(ctx, delta, map_, sensor, T0):
            d = delta.detach().cpu().double().numpy()
            poses = np.stack([se3_exp(d[k]) @ T0[k] for k in range(len(T0))])
            depth, grad, valid = raycast_depth_grad(map_, sensor, poses)
            ctx.save_for_backward(grad)
            ctx.mark_non_differentiable(valid)
            return depth.to(delta.dtype), valid

        @staticmethod
        def backward(ctx, g_depth, g_valid):
            (grad,) = ctx.saved_tensors
            return (g_depth[..., None] * grad).sum(dim=(1, 2)).to(g_depth.dtype), None, None, None

    return PointToPlaneCost, RaycastDepth


def point_to_plane_cost(delta, tracker, depth, ref_points, ref_normals, T_ref, T0):
    return _autograd()[0].apply(delta, tracker, depth, ref_points, ref_normals, T_ref, T0)


def raycast_depth(delta, map_, sensor, T0):
    return _autograd()[1].apply(delta, map_, sensor, T0)

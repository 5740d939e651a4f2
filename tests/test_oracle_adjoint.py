"""FP64 oracle of the differentiable raycast depth: the implicit-function-theorem gradient against
central finite differences of the refined root (eps 1e-5), and the refined depth against the
reference march's linear zero crossing."""
import numpy as np

from oracle import oracle as O


def _tilted_wall_map(pin, precision="f64"):
    H, W = pin.height, pin.width
    v, u = np.mgrid[0:H, 0:W].astype(np.float64)
    dx = (u - pin.cx) / pin.fx
    depth = (2.0 / (1 - 0.2 * dx)).astype(np.float32)  # the plane z = 2 + 0.2 x seen from the origin
    m = O.OracleMap(10.24, 0.01, precision=precision)
    for k in range(2):
        m.integrate(k, depth, pin, np.eye(4))
    return m


def test_depth_gradient_matches_central_differences():
    pin = O.pinhole(160, 120, 131.25, 131.25, 79.5, 59.5, 0.4, 6.0)
    m = _tilted_wall_map(pin)
    T0 = np.eye(4)
    T0[:3, 3] = [0.003, -0.002, 0.001]
    d0, g0, v0 = m.raycast_depth_grad(pin, T0)
    assert v0.mean() > 0.85
    eps = 1e-5
    fd = np.zeros_like(g0)
    ok = v0.copy()
    for k in range(6):
        xi = np.zeros(6)
        xi[k] = eps
        dp, _, vp = m.raycast_depth_grad(pin, O.se3_exp(xi, "f64") @ T0)
        dm, _, vm = m.raycast_depth_grad(pin, O.se3_exp(-xi, "f64") @ T0)
        fd[..., k] = (dp - dm) / (2 * eps)
        ok &= vp & vm
    err = np.abs(fd - g0)[ok]
    # exact inside a trilinear cell; rays whose root crosses a cell face within +-eps see a kink
    assert np.median(err) < 1e-8
    assert np.mean(err < 1e-5) > 0.99
    # moving the camera along its optical axis changes depth about one-for-one at the image centre
    # (the projective TSDF of a tilted plane is not exactly the plane's distance field)
    assert abs(g0[59, 79, 2] + 1.0) < 5e-3


def test_refined_depth_close_to_linear_crossing():
    pin = O.pinhole(160, 120, 131.25, 131.25, 79.5, 59.5, 0.4, 6.0)
    m = _tilted_wall_map(pin, precision="f32")
    T = np.eye(4, dtype=np.float32)
    d, _, v = m.raycast_depth_grad(pin, T)
    pts, _, _ = m.raycast(pin, T)
    both = v & np.any(pts != 0, -1)
    assert both.mean() > 0.8
    assert np.abs(d[both] - pts[both][:, 2]).max() < 0.01

"""Precision variant: FP16 TSDF + uint16 weights against the FP32 path and the FP64 oracle.

Bar: allocation bit-exact with the FP32 path; TSDF RMS error vs the FP64
oracle <= 2e-4 m (error in metres = |dtsdf| * mu); weights identical."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import CAM, circle_frames

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MU = 0.08


def _sensor(se):
    return se.PinholeCamera(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)


def _pin():
    return O.pinhole(CAM.width, CAM.height, CAM.fx, CAM.fy, CAM.cx, CAM.cy, CAM.near_plane, CAM.far_plane)


def _common(k1, k2):
    common, i1, i2 = np.intersect1d(k1, k2, assume_unique=True, return_indices=True)
    return common, i1, i2


def test_fp16_tsdf_rms_vs_fp64_oracle():
    from supereight2_amd import se

    n = 8
    depth, poses = circle_frames(50, n)
    sensor = _sensor(se)
    maps = {s: se.TSDFMap(10.24, 0.01, storage=s) for s in ("f32", "f16")}
    o64 = O.OracleMap(10.24, 0.01, precision="f64")
    pin = _pin()
    for k in range(n):
        print(f"frame {k}", flush=True)
        d = torch.from_numpy(depth[k]).cuda()
        for m in maps.values():
            se.MapIntegrator(m).integrateDepth(k, d, sensor, poses[k])
        o64.integrate(k, depth[k], pin, poses[k].astype(np.float64))
    k32, _, t32, w32 = maps["f32"].dump_blocks()
    k16, _, t16, w16 = maps["f16"].dump_blocks()
    np.testing.assert_array_equal(k16, k32)  # allocation does not depend on the storage type
    np.testing.assert_array_equal(w16, w32)
    k64, _, t64, w64 = o64.dump_blocks()
    common, i32, i64 = _common(k32, k64)
    assert len(common) > 0.99 * len(k32)
    both = (w32[i32] > 0) & (w64[i64] > 0)
    e32 = (t32[i32] - t64[i64])[both] * MU
    e16 = (t16[i32] - t64[i64])[both] * MU
    rms32 = float(np.sqrt(np.mean(e32.astype(np.float64) ** 2)))
    rms16 = float(np.sqrt(np.mean(e16.astype(np.float64) ** 2)))
    print(f"TSDF RMS vs FP64 oracle: fp32 {rms32:.3e} m, fp16 {rms16:.3e} m over {both.sum()} voxels")
    assert rms16 <= 2e-4
    assert rms32 <= rms16
    # fp16 rounding bound against the fp32 path, whose update history (allocation, band test,
    # weights) is the same voxel for voxel: at most half an fp16 ulp of |tsdf| <= 1 (2^-12) per
    # update, shrunk by w/(w+1) at the next one, so below n * 2^-12 after n frames.  (Against the
    # FP64 oracle a voxel on the edge of the band can take one more or one fewer update, so the
    # bound is stated against fp32; the FP64 comparison is the RMS above.)
    upd = w32 > 0
    assert np.abs(t16 - t32)[upd].max() < n * 2.0**-12

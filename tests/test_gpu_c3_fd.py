"""Geometry gradients at the 256^3 x 720-view workload against FP64 finite differences (GPU).

For rays of one view of the perturbed 256^3 scan (768x640 detector, pixel 0.6 mm, SOD 600,
SID 1100, voxel 0.5, step 0.25), the GPU adjoint's eight geometry derivatives of one ray's
projection (a one-hot upstream gradient) are compared with central differences of the FP64
oracle's ray integral.  The projection is piecewise smooth -- it jumps where a sample changes
interpolation cell, where the sample count changes or where the entry face changes -- so, as
in the reference's own check (examples/check_forward_operator.py:179-192), a ray is used only
where its sampling trace (entry face, sample count, hash of the visited cells) is the same at
both finite-difference points as at the base point, for every parameter, and where the FP32
schedule the GPU follows visits the same cells as the FP64 one.

Tolerance: the reference's for its GPU gradients (rel 1e-3 / abs 1e-4,
check_forward_operator.py:226), set on its 2-view fixture.  A ray of this scan sums ~400
samples whose angle terms carry a 600 mm lever arm and cancel to a small total, so an FP32
evaluation of the reference's own algorithm can miss that tolerance on some rays: there the
GPU must be no farther from the finite difference than twice the distance of the FP32
restatement of the reference kernel (the oracle's FP32 schedule, same one-hot upstream).  Every
comparison must also meet the reference's bound for FP32 GPU finite differences (rel 1e-2 /
abs 2e-3, check_forward_operator.py:256).

To keep the oracle's work small the detector is cropped to a 16x16 window: pixel u of the
window is pixel u + u_lo of the full detector when det_u0 is lowered by u_lo (likewise v), so
the rays and every derivative, det_u0's and det_v0's included, are those of the full scan."""

import numpy as np
import pytest
import torch

import bench
import oracle
from cbct_backprojector import _native

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
# finite-difference steps per parameter: angle (rad), sid, sod (mm), det_u0, det_v0 (px), src (mm)
EPS = (1e-7, 1e-5, 1e-5, 1e-5, 1e-5, 1e-5, 1e-5, 1e-5)
NAMES = ("angle", "sid", "sod", "det_u0", "det_v0", "src_x", "src_y", "src_z")


def _crop(geom_np, view, u_lo, v_lo):
    g = [np.asarray(x[view:view + 1], np.float32).copy() for x in geom_np]
    g[3] = g[3] - np.float32(u_lo)
    g[4] = g[4] - np.float32(v_lo)
    return g


@pytest.mark.timeout(600)
@pytest.mark.parametrize("view,du,dv", [(250, 40, -30), (611, -60, 25)])
def test_c3_geometry_gradients_match_fp64_finite_differences(view, du, dv):
    w = bench.WORKLOADS["geometry-gradients-256c-720v"]
    vol_np, geom_np, vol, _, _ = bench.make_inputs(w, DEV)
    n = 16
    u_lo, v_lo = w["U"] // 2 - n // 2 + du, w["V"] // 2 - n // 2 + dv
    g32 = _crop(geom_np, view, u_lo, v_lo)
    pix, step, vs = w["pixel"], w["step"], w["voxel"]
    vol64 = vol_np.astype(np.float64)
    g64 = [x.astype(np.float64) for x in g32]
    _, tr64 = oracle.fp_forward(vol64, g64, pix, n, n, step, vs, mode="f64", traces=True)
    _, tr32 = oracle.fp_forward(vol_np, g32, pix, n, n, step, vs, mode="f32", traces=True)
    same = lambda a, b: (a[0] == b[0]) & (a[1] == b[1]) & (a[2] == b[2])
    stable = same(tr32, tr64)[0] & (tr64[1][0] > 0)
    fd = np.zeros((8, n, n))
    for k in range(8):
        plus = [x.copy() for x in g64]
        minus = [x.copy() for x in g64]
        plus[k] += EPS[k]
        minus[k] -= EPS[k]
        pp, trp = oracle.fp_forward(vol64, plus, pix, n, n, step, vs, mode="f64", traces=True)
        pm, trm = oracle.fp_forward(vol64, minus, pix, n, n, step, vs, mode="f64", traces=True)
        stable &= same(trp, tr64)[0] & same(trm, tr64)[0]
        fd[k] = (pp[0] - pm[0]) / (2.0 * EPS[k])
    rays = [(u, v) for u in range(n) for v in range(n) if stable[u, v]]
    print(f"view {view}, window ({u_lo}, {v_lo}): {len(rays)} of {n * n} rays trace-stable for all 8 parameters",
          flush=True)
    assert len(rays) >= 8
    rays = rays[:: max(1, len(rays) // 8)][:8]
    gdev = [torch.tensor(x, device=DEV) for x in g32]
    strict, worst = 0, 0.0
    for (u, v) in rays:
        up = torch.zeros((1, n, n), device=DEV)
        up[0, u, v] = 1.0
        _, gg = _native.fp_backward(up, vol, gdev, pix, step, vs, False, 0., 0., 0.)
        got = gg[:, 0].double().cpu().numpy()
        _, g_f32 = oracle.fp_backward(up.cpu().numpy(), vol_np, g32, pix, step, vs, need_volume=False, mode="f32")
        ref = fd[:, u, v]
        err = np.abs(got - ref)
        err_f32 = np.abs(g_f32[:, 0].astype(np.float64) - ref)
        tol = 1e-4 + 1e-3 * np.abs(ref)
        strict += int((err <= tol).sum())
        worst = max(worst, float((err / tol).max()))
        for k in range(8):
            assert err[k] <= tol[k] or err[k] <= 2.0 * err_f32[k], \
                (u, v, NAMES[k], got[k], ref[k], g_f32[k, 0])
            assert err[k] <= 2e-3 + 1e-2 * abs(ref[k]), (u, v, NAMES[k], got[k], ref[k])
    print(f"  8 rays x 8 parameters: {strict} of 64 within rel 1e-3 / abs 1e-4 of the FD, the rest within twice the "
          f"FP32 schedule's own distance; largest |GPU - FD| / (1e-4 + 1e-3 |FD|) = {worst:.3f}", flush=True)

"""Parity at the benchmark sizes (GPU).

Views are independent in the forward projection and in the per-view geometry
gradients, so the full-size workloads are checked on a bounded subset of views
against the CPU oracle, and as a whole through size-independent properties
(constant volume -> exact sample counts, linearity).  The FDK reconstruction is
checked on a slab of slices against the FP64 oracle.
"""

import math

import numpy as np
import pytest
import torch

import bench
import cases
import oracle
from cbct_backprojector import _native

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


@pytest.fixture(scope="module")
def c3():
    w = bench.WORKLOADS["geometry-gradients-256c-720v"]
    vol_np, geom_np, vol, geom, up = bench.make_inputs(w, DEV)
    return w, vol_np, geom_np, vol, geom, up


def _subset(geom_np, geom, views):
    return [g[views].copy() for g in geom_np], [g[views].contiguous() for g in geom]


def test_c3_forward_subset_matches_fp32_schedule(c3):
    w, vol_np, geom_np, vol, geom, _ = c3
    proj = _native.fp_forward(vol, geom, w["pixel"], w["U"], w["V"], w["step"], w["voxel"], 0., 0., 0.)
    views = np.array([0, 181, 362, 543])
    g_np, _ = _subset(geom_np, geom, views)
    args = (w["pixel"], w["U"], w["V"], w["step"], w["voxel"])
    ref32 = oracle.fp_forward(vol_np, g_np, *args, mode="f32")
    got = cases.to_np(proj)[views]
    e32 = cases.rel_l2(got, ref32)
    ref64 = oracle.fp_forward(vol_np.astype(np.float64), g_np, *args, mode="f64")
    e64 = cases.rel_l2(got, ref64)
    print(f"256^3 x 720 forward, 4 views: rel L2 vs FP32 schedule {e32:.3e}, vs FP64 {e64:.3e}")
    assert e32 <= 1e-6 and e64 <= 1e-5


def test_c3_adjoint_subset_matches_oracle(c3):
    """Volume gradient and geometry gradients of a 4-view sub-scan at full detector/volume size.
    The volume gradient is compared with the FP32 oracle.  Each geometry gradient is a sum over
    ~1e8 samples with cancellation, so two FP32 evaluation orders differ by ~1e-5: the GPU's
    sums are compared with the FP64 oracle and must be at least as close to it as the FP32
    oracle's are (or within 1e-5)."""
    w, vol_np, geom_np, vol, geom, _ = c3
    views = np.array([7, 250, 489, 700])
    g_np, g_dev = _subset(geom_np, geom, views)
    rng = np.random.default_rng(21)
    up = rng.standard_normal((len(views), w["U"], w["V"])).astype(np.float32)
    gv, gg = _native.fp_backward(torch.tensor(up, device=DEV), vol, g_dev, w["pixel"], w["step"], w["voxel"],
                                 True, 0., 0., 0.)
    rvol, rgeo = oracle.fp_backward(up, vol_np, g_np, w["pixel"], w["step"], w["voxel"], mode="f32")
    _, rgeo64 = oracle.fp_backward(up.astype(np.float64), vol_np.astype(np.float64), g_np, w["pixel"],
                                   w["step"], w["voxel"], mode="f64")
    e_vol = cases.rel_l2(cases.to_np(gv), rvol)
    print(f"256^3 x 4-view adjoint: volume gradient rel L2 vs FP32 oracle {e_vol:.3e}")
    assert e_vol <= 1e-5
    for k in range(8):
        e32 = cases.rel_l2(cases.to_np(gg[k]), rgeo[k])
        e_gpu = cases.rel_l2(cases.to_np(gg[k]), rgeo64[k])
        e_f32 = cases.rel_l2(rgeo[k], rgeo64[k])
        print(f"  geometry gradient {k}: GPU vs FP32 oracle {e32:.2e}, GPU vs FP64 {e_gpu:.2e}, "
              f"FP32 oracle vs FP64 {e_f32:.2e}")
        assert e_gpu <= max(1e-5, 2.0 * e_f32)


def test_c3_constant_volume_gives_exact_sample_counts(c3):
    """proj of a constant volume = count * step per ray, so sum(proj)/step = S exactly (to FP32 rounding)."""
    w, _, _, vol, geom, _ = c3
    ones = torch.ones_like(vol)
    proj = _native.fp_forward(ones, geom, w["pixel"], w["U"], w["V"], w["step"], w["voxel"], 0., 0., 0.)
    S = _native.fp_count_samples(tuple(vol.shape), geom, w["pixel"], w["U"], w["V"], w["step"], w["voxel"])
    total = proj.double().sum().item() / w["step"]
    assert abs(total - S) / S <= 1e-6
    per_ray = torch.round(proj.double() / w["step"])
    assert torch.allclose(proj.double(), per_ray * w["step"], rtol=2e-6, atol=1e-6)


def test_c3_linearity_full_size(c3):
    w, _, _, vol, geom, _ = c3
    other = vol.flip(-1).contiguous()
    args = (w["pixel"], w["U"], w["V"], w["step"], w["voxel"], 0., 0., 0.)
    p1 = _native.fp_forward(vol, geom, *args)
    p2 = _native.fp_forward(other, geom, *args)
    p3 = _native.fp_forward((1.3 * vol - 0.4 * other).contiguous(), geom, *args)
    err = ((p3 - (1.3 * p1 - 0.4 * p2)).double().norm() / p3.double().norm()).item()
    print(f"256^3 x 720 linearity rel L2 {err:.3e}")
    assert err <= 1e-5     # FP32 rounding of the combined volume; the reference allows 5e-5


def test_c2_fdk_roundtrip_slab_matches_fp64_oracle():
    """Project the 256^3 phantom over 360 views, reconstruct with FDK, compare 8 slices
    with the FP64 oracle (filter in FP64, chunked weighted backprojection)."""
    w = bench.WORKLOADS["fdk-roundtrip-256c-360v"]
    vol_np, geom_np, vol, geom, _ = bench.make_inputs(w, DEV)
    D = w["D"]
    proj = _native.fp_forward(vol, geom, w["pixel"], w["U"], w["V"], w["step"], w["voxel"], 0., 0., 0.)
    rec = _native.fdk_reconstruct(proj, geom, w["pixel"], D, D, D, w["voxel"], 0., 0., 0., 0.8, 32)
    p_np = cases.to_np(proj).astype(np.float64)
    filt = oracle.fdk_filter(p_np, w["pixel"], geom_np[2], geom_np[1], 0.8)
    z0, nz = 120, 8
    offz = (z0 - D / 2 + nz / 2) * w["voxel"]
    slab = oracle.bp_forward(filt, geom_np, w["pixel"], (nz, D, D), w["voxel"], (0.0, 0.0, offz), chunk=32,
                             mode="f64")
    slab = np.maximum(slab * (math.pi / w["A"]), 0.0)
    got = cases.to_np(rec)[z0:z0 + nz]
    err = cases.rel_l2(got, slab)
    acc = cases.rel_l2(got, vol_np[z0:z0 + nz])
    print(f"FDK 256^3 x 360: slab rel L2 vs FP64 oracle {err:.3e}; vs phantom {acc:.3f}")
    assert err <= 1e-5
    assert acc < 0.5


def test_c2_backprojection_backward_subset_matches_oracle():
    """The backprojection backward (LDS detector windows, fixed-point accumulation) at the
    256^3 x 512x384 scale, on 3 views.  The projection gradient is compared with the FP32
    oracle.  Each geometry gradient is a sum over 16.7M voxels of random-sign terms, so two FP32
    summation orders differ by ~1e-5; the GPU's sums are compared with the FP64 oracle and must
    be at least as close to it as the FP32 oracle's (or within 1e-5)."""
    w = bench.WORKLOADS["backprojection-256c-360v"]
    vol_np, geom_np, vol, geom, _ = bench.make_inputs(w, DEV)
    views = np.array([5, 170, 301])
    g_np, g_dev = _subset(geom_np, geom, views)
    proj = _native.fp_forward(vol, g_dev, w["pixel"], w["U"], w["V"], w["step"], w["voxel"], 0., 0., 0.)
    rng = np.random.default_rng(33)
    D = w["D"]
    gv = rng.standard_normal((D, D, D)).astype(np.float32)
    gproj, ggeo = _native.bp_backward(torch.tensor(gv, device=DEV), proj, g_dev, w["pixel"], w["voxel"], 0., 0., 0.)
    p_np = cases.to_np(proj)
    rproj, rgeo32 = oracle.bp_backward(gv, p_np, g_np, w["pixel"], w["voxel"], mode="f32")
    _, rgeo64 = oracle.bp_backward(gv.astype(np.float64), p_np.astype(np.float64), g_np, w["pixel"], w["voxel"],
                                   mode="f64")
    e_proj = cases.rel_l2(cases.to_np(gproj), rproj)
    print(f"256^3 backprojection backward, 3 views: projection gradient vs FP32 oracle {e_proj:.2e}")
    assert e_proj <= 1e-5
    for k in range(8):
        e_gpu = cases.rel_l2(cases.to_np(ggeo[k]), rgeo64[k])
        e_f32 = cases.rel_l2(rgeo32[k], rgeo64[k])
        print(f"  geometry gradient {k}: GPU vs FP64 {e_gpu:.2e}, FP32 oracle vs FP64 {e_f32:.2e}")
        assert e_gpu <= max(1e-5, 2.0 * e_f32)


def test_fused_calibration_gradient_at_medium_size():
    """The fused calibration march against the plain operators on a 128^3 x 24-view problem with
    a 256^2 detector (the residual scale 2/(A U V) and the per-tile partials at realistic size)."""
    from cbct_backprojector import calibration
    spec = {"D": 128, "A": 24, "U": 256, "V": 256, "voxel": 0.5, "sod": 600.0, "sid": 1100.0, "seed": 77,
            "pixel": (1100 / 600) * 1.42 * 128 * 0.5 / (0.95 * 256)}
    p = calibration.synthetic_problem(spec, DEV)
    g0 = p.geometry.clone()
    calibration.calibrate([p], steps=0, want_grad=True)
    A, U, V = p.measured.shape
    proj = _native.fp_forward(p.volume, list(g0), p.pixel_size, U, V, p.step_size, p.voxel_size, 0., 0., 0.)
    r = proj - p.measured
    _, gg = _native.fp_backward(((2.0 / (A * U * V)) * r).contiguous(), p.volume, list(g0), p.pixel_size,
                                p.step_size, p.voxel_size, False, 0., 0., 0.)
    loss = (r.double() ** 2).mean().item()
    assert abs(p.loss_history[0].item() - loss) <= 1e-5 * loss
    errs = [cases.rel_l2(cases.to_np(p.grad[k]), cases.to_np(gg[k])) for k in range(8)]
    print("fused vs plain at 128^3 x 24 x 256^2:", ["%.1e" % e for e in errs])
    assert max(errs) <= 1e-4

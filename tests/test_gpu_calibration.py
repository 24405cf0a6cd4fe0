"""Batched geometry calibration (GPU).

A fused calibration step (march -> residual -> geometry gradient) must give the
same loss and gradient as the plain operators (forward projection, residual,
adjoint with the residual as upstream gradient); steps of several uneven
problems run together must equal each problem run alone; the Adam trajectory
must follow an FP64 CPU restatement of Adam driven by the plain operators; and
the fit must reduce the reprojection error.
"""

import numpy as np
import pytest
import torch

import oracle
from cbct_backprojector import _native, calibration

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")

SMALL = {"D": 40, "A": 18, "U": 70, "V": 50, "voxel": 0.5, "sod": 600.0, "sid": 1100.0, "seed": 5}
SMALL["pixel"] = (1100 / 600) * 1.42 * 40 * 0.5 / (0.95 * 50)


def _problem(spec, **kw):
    return calibration.synthetic_problem(spec, DEV, **kw)


def _plain_loss_grad(p, geom):
    """Loss and geometry gradient through the plain operators."""
    glist = [g.contiguous() for g in geom]
    A, U, V = p.measured.shape
    proj = _native.fp_forward(p.volume, glist, p.pixel_size, U, V, p.step_size, p.voxel_size, 0., 0., 0.)
    r = proj - p.measured
    loss = (r.double() ** 2).mean().item()
    up = (2.0 / (A * U * V)) * r
    _, gg = _native.fp_backward(up.contiguous(), p.volume, glist, p.pixel_size, p.step_size, p.voxel_size,
                                False, 0., 0., 0.)
    return loss, torch.stack(list(gg))


def _rel(a, b):
    return ((a.double() - b.double()).norm() / b.double().norm().clamp_min(1e-30)).item()


@pytest.mark.parametrize("use_graph", [False, True])
def test_fused_gradient_equals_plain_operators(use_graph):
    """Zero steps: the loss and its geometry gradient at the start geometry.  The fused march
    weights each ray's geometry sums by its residual once at the end, the plain adjoint weights
    every sample; the angle gradient sums terms of both signs over the view, so the two FP32
    orders are compared through their distance to the FP64 oracle."""
    p = _problem(SMALL)
    g0 = p.geometry.clone()
    calibration.calibrate([p], steps=0, use_graph=use_graph, want_grad=True)
    assert torch.equal(p.geometry, g0)                      # no update at zero steps
    loss, grad = _plain_loss_grad(p, g0)
    e_loss = abs(p.loss_history[0].item() - loss) / loss
    A, U, V = p.measured.shape
    proj = _native.fp_forward(p.volume, list(g0), p.pixel_size, U, V, p.step_size, p.voxel_size, 0., 0., 0.)
    up = ((2.0 / (A * U * V)) * (proj - p.measured)).cpu().numpy()
    _, ref64 = oracle.fp_backward(up, p.volume.cpu().numpy(), g0.cpu().numpy(), p.pixel_size, p.step_size,
                                  p.voxel_size, need_volume=False, mode="f64")
    ref64 = torch.tensor(ref64)
    fused, plain = p.grad.cpu(), grad.cpu()
    e_fp = [_rel(fused[k], plain[k]) for k in range(8)]
    e_f64 = [_rel(fused[k], ref64[k]) for k in range(8)]
    e_p64 = [_rel(plain[k], ref64[k]) for k in range(8)]
    fmt = lambda e: ["%.1e" % x for x in e]
    print(f"loss rel {e_loss:.2e}; fused vs plain {fmt(e_fp)}; fused vs FP64 {fmt(e_f64)}; plain vs FP64 {fmt(e_p64)}")
    assert e_loss <= 1e-5
    for k in range(8):
        assert e_fp[k] <= 1e-5 or e_f64[k] <= 2.0 * e_p64[k]
        assert e_fp[k] <= 1e-4


def test_batched_problems_equal_individual_runs():
    specs = [dict(SMALL), dict(SMALL, D=33, A=11, U=40, V=40, seed=9), dict(SMALL, D=52, A=25, U=90, V=66, seed=13)]
    batch = [_problem(s) for s in specs]
    calibration.calibrate(batch, steps=3, use_graph=True)
    for s, pb in zip(specs, batch):
        pa = _problem(s)
        calibration.calibrate([pa], steps=3, use_graph=False)
        assert torch.equal(pa.geometry, pb.geometry)
        assert torch.allclose(pa.loss_history, pb.loss_history, rtol=1e-6, atol=0)


def test_adam_trajectory_matches_fp64_adam_on_plain_operators():
    steps, lr, b1, b2, eps = 4, calibration.DEFAULT_LR, 0.9, 0.999, 1e-8
    p = _problem(SMALL)
    g = p.geometry.clone()
    calibration.calibrate([p], steps=steps, use_graph=True)
    # FP64 Adam on the CPU, gradients from the plain GPU operators
    geo = g.double().cpu().numpy()
    m = np.zeros_like(geo)
    v = np.zeros_like(geo)
    lr_col = np.asarray(lr, dtype=np.float64)[:, None]
    losses = []
    for t in range(1, steps + 1):
        loss, grad = _plain_loss_grad(p, torch.tensor(geo, dtype=torch.float32, device=DEV))
        losses.append(loss)
        gr = grad.double().cpu().numpy()
        m = b1 * m + (1 - b1) * gr
        v = b2 * v + (1 - b2) * gr * gr
        geo = geo - lr_col * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    got = p.geometry.double().cpu().numpy()
    dev = np.abs(got - geo).max(axis=1) / np.asarray(lr)
    print("max |fused - FP64 Adam| per group, in units of lr:", np.array2string(dev, precision=2))
    assert dev.max() <= 0.05
    hist = p.loss_history.double().cpu().numpy()
    assert np.allclose(hist[:steps], losses, rtol=1e-4)


def test_calibration_reduces_reprojection_error():
    spec = dict(SMALL, D=48, A=24, U=64, V=64, seed=21)
    spec["pixel"] = (1100 / 600) * 1.42 * 48 * 0.5 / (0.95 * 64)
    p = _problem(spec)
    shape = tuple(p.volume.shape)
    e0 = calibration.reprojection_error_px(p.geometry, p.true_geometry, shape, p.voxel_size, p.pixel_size)
    calibration.calibrate([p], steps=40, use_graph=True)
    e1 = calibration.reprojection_error_px(p.geometry, p.true_geometry, shape, p.voxel_size, p.pixel_size)
    hist = p.loss_history.cpu().numpy()
    print(f"40 steps: loss {hist[0]:.3e} -> {hist[-1]:.3e}; reprojection error {e0:.3f} px -> {e1:.3f} px")
    assert hist[-1] < 0.2 * hist[0]
    assert e1 < 0.5 * e0


def test_continued_fit_equals_one_long_fit():
    """calibrate(a steps) then calibrate(b steps, first_step=a) with the same Adam moments
    equals calibrate(a + b steps), learning-rate decay included."""
    p1 = _problem(SMALL)
    calibration.calibrate([p1], steps=5, use_graph=True, lr_decay=0.9)
    p2 = _problem(SMALL)
    calibration.calibrate([p2], steps=2, use_graph=False, lr_decay=0.9)
    calibration.calibrate([p2], steps=3, use_graph=True, lr_decay=0.9, first_step=2)
    assert torch.equal(p1.geometry, p2.geometry)
    assert torch.equal(p1.adam_m, p2.adam_m) and torch.equal(p1.adam_v, p2.adam_v)

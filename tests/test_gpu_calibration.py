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
# a voxel size that is not a power of two: the fused march takes the forward's division map
SMALL_DIV = dict(SMALL, voxel=0.8, seed=6)
SMALL_DIV["pixel"] = (1100 / 600) * 1.42 * 40 * 0.8 / (0.95 * 50)


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


@pytest.mark.parametrize("use_graph,spec", [(False, SMALL), (True, SMALL), (False, SMALL_DIV)])
def test_fused_gradient_equals_plain_operators(use_graph, spec):
    """Zero steps: the loss and its geometry gradient at the start geometry.  The fused march
    weights each ray's geometry sums by its residual once at the end, the plain adjoint weights
    every sample; the angle gradient sums terms of both signs over the view, so the two FP32
    orders are compared through their distance to the FP64 oracle."""
    p = _problem(spec)
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


def test_lm_step_solves_the_gauss_newton_normal_equations():
    """One Levenberg-Marquardt iteration with negligible damping: the geometry change delta
    satisfies H delta = -g, with H = sum j j^T and g = sum j r built in FP64 from per-ray
    derivatives j of the plain adjoint (one-hot upstream gradients) and residuals r of the
    plain forward projection."""
    spec = dict(SMALL, D=24, A=2, U=8, V=8, seed=17)
    spec["pixel"] = (1100 / 600) * 1.42 * 24 * 0.5 / (0.95 * 8)
    p = _problem(spec)
    g0 = p.geometry.clone()
    A, U, V = p.measured.shape
    glist = [g.contiguous() for g in g0]
    proj = _native.fp_forward(p.volume, glist, p.pixel_size, U, V, p.step_size, p.voxel_size, 0., 0., 0.)
    r = (proj - p.measured).double().cpu()
    J = torch.zeros((A, U, V, 8), dtype=torch.float64)
    for u in range(U):
        for v in range(V):
            up = torch.zeros((A, U, V), device=DEV)
            up[:, u, v] = 1.0
            _, gg = _native.fp_backward(up, p.volume, glist, p.pixel_size, p.step_size, p.voxel_size, False,
                                        0., 0., 0.)
            J[:, u, v, :] = torch.stack(list(gg)).double().cpu().T
    calibration.calibrate([p], steps=1, solver="lm", lm_lambda=1e-6, use_graph=False)
    delta = (p.geometry.double() - g0.double()).cpu()
    assert delta.abs().max() > 0, "the first iteration was rejected"
    for a in range(A):
        Ja = J[a].reshape(-1, 8)
        H = Ja.T @ Ja
        g = Ja.T @ r[a].reshape(-1)
        res = (H @ delta[:, a] + g).norm() / g.norm()
        print(f"view {a}: |H delta + g| / |g| = {res:.2e}")
        assert res <= 1e-2


def test_lm_converges_monotonically_and_beats_adam_per_iteration():
    spec = dict(SMALL, D=48, A=24, U=64, V=64, seed=21)
    spec["pixel"] = (1100 / 600) * 1.42 * 48 * 0.5 / (0.95 * 64)
    p_lm, p_adam = _problem(spec), _problem(spec)
    shape = tuple(p_lm.volume.shape)
    calibration.calibrate([p_lm], steps=10, solver="lm", use_graph=True)
    calibration.calibrate([p_adam], steps=10, use_graph=True)
    hist = p_lm.loss_history.cpu().numpy()
    assert np.all(np.diff(hist) <= 0)
    e_lm = calibration.reprojection_error_px(p_lm.geometry, p_lm.true_geometry, shape, p_lm.voxel_size,
                                             p_lm.pixel_size)
    e_adam = calibration.reprojection_error_px(p_adam.geometry, p_adam.true_geometry, shape, p_adam.voxel_size,
                                               p_adam.pixel_size)
    print(f"10 iterations: LM {e_lm:.3f} px (loss {hist[0]:.2e} -> {hist[-1]:.2e}), Adam {e_adam:.3f} px")
    assert e_lm < e_adam and e_lm < 0.2


def test_lm_batched_problems_equal_individual_runs():
    specs = [dict(SMALL), dict(SMALL, D=33, A=11, U=40, V=40, seed=9)]
    batch = [_problem(s) for s in specs]
    calibration.calibrate(batch, steps=3, solver="lm", use_graph=True)
    for s, pb in zip(specs, batch):
        pa = _problem(s)
        calibration.calibrate([pa], steps=3, solver="lm", use_graph=False)
        assert torch.equal(pa.geometry, pb.geometry)


def test_autograd_loop_with_detector_shifts_follows_the_fused_adam():
    """The plain calibration loop through the public API (leaf tensors, detector_centers for the
    detector's physical shifts, project, backward, torch Adam) and the fused batched Adam on the
    centre indices follow the same trajectory when the centre's learning rate is the shift's
    divided by the pixel size (Adam normalises each gradient by its own RMS)."""
    steps = 4
    p = _problem(SMALL)
    g0 = p.geometry.clone()
    A, U, V = p.measured.shape
    lr_shift = list(calibration.DEFAULT_LR)
    lr_shift[3] *= p.pixel_size
    lr_shift[4] *= p.pixel_size
    geo_ag, shifts, hist_ag = calibration.calibrate_autograd(p.volume, p.measured, g0, p.pixel_size, p.step_size,
                                                             p.voxel_size, steps=steps, lr=lr_shift)
    calibration.calibrate([p], steps=steps, use_graph=True)
    dev = ((geo_ag - p.geometry).abs().max(dim=1).values.cpu().double().numpy()
           / np.asarray(calibration.DEFAULT_LR))
    print("max |autograd loop - fused Adam| per group, in units of lr:", np.array2string(dev, precision=3))
    assert dev.max() <= 0.05
    u0 = (U - 1) / 2.0 - shifts["delta_u"] / p.pixel_size
    assert torch.allclose(u0, geo_ag[3], rtol=0, atol=1e-4)
    assert torch.allclose(hist_ag.float(), p.loss_history, rtol=1e-3)


def test_fitting_only_the_detector_shifts_recovers_them():
    """A scan whose only miscalibration is a per-view detector shift: fitting delta_u / delta_v
    alone through detector_centers by autograd recovers it."""
    spec = dict(SMALL, seed=31)
    p = _problem(spec, perturb=False)
    A, U, V = p.measured.shape
    true = p.geometry.clone()
    gen = torch.Generator(device="cpu").manual_seed(3)
    shift = (torch.randn(2, A, generator=gen) * 0.4).to(DEV)          # ~0.4 mm
    true[3] -= shift[0] / p.pixel_size
    true[4] -= shift[1] / p.pixel_size
    measured = _native.fp_forward(p.volume, list(true), p.pixel_size, U, V, p.step_size, p.voxel_size, 0., 0., 0.)
    lr = [0.0] * 8
    lr[3] = lr[4] = 0.05 * p.pixel_size * 4
    geo, shifts, hist = calibration.calibrate_autograd(p.volume, measured, p.geometry, p.pixel_size, p.step_size,
                                                       p.voxel_size, steps=60, lr=lr, fit=("delta_u", "delta_v"))
    err0 = (shift[0].abs().mean() + shift[1].abs().mean()).item() / 2
    err1 = ((shifts["delta_u"] - shift[0]).abs().mean() + (shifts["delta_v"] - shift[1]).abs().mean()).item() / 2
    print(f"detector shift error {err0:.3f} mm -> {err1:.3f} mm; loss {hist[0].item():.2e} -> {hist[-1].item():.2e}")
    assert err1 < 0.2 * err0
    assert torch.equal(geo[0], p.geometry[0]) and torch.equal(geo[5], p.geometry[5])


def test_graph_cache_eviction_reuse_and_release():
    """More distinct graphs than the cache keeps (eight): keys captured again after their eviction
    give the same fit bit for bit, the release frees the cached and the evicted graphs, and a call
    after it captures afresh with the same result."""
    runs = {}
    for steps in list(range(1, 11)) + [1, 2]:
        p = _problem(SMALL)
        calibration.calibrate([p], steps=steps, use_graph=True)
        runs.setdefault(steps, []).append((p.geometry.clone(), p.loss_history.clone()))
    for steps in (1, 2):
        (g0, l0), (g1, l1) = runs[steps]
        assert torch.equal(g0, g1) and torch.equal(l0, l1)
    torch.cuda.synchronize()
    assert _native.lib().cbct_calib_release_graphs() == 0
    p = _problem(SMALL)
    calibration.calibrate([p], steps=3, use_graph=True)
    g3, l3 = runs[3][0]
    assert torch.equal(p.geometry, g3) and torch.equal(p.loss_history, l3)

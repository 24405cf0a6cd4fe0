"""Batched calibration at its full workload (GPU): 64 problems of uneven size (cubic ellipsoid
phantoms of side 96-320, 90-720 views, square detectors 128-768, voxel 0.5 mm, SOD 600, SID
1100, geometry perturbed per view), 20 Adam steps each, all in one hipGraph of persistent
launches -- the same problem set as the benchmark's batched-calibration workload.

Checks: the step-0 loss and geometry gradient of every problem equal those of the plain
operators (forward projection, residual, geometry adjoint with the residual as upstream
weight); the smallest, the median and the largest problem fitted alone equal their batched fit
bit for bit; the fit lowers every problem's loss and the median reprojection error."""

import numpy as np
import pytest
import torch

from cbct_backprojector import _native, calibration, phantom

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
STEPS = 20


def _plain_loss_grad(p, geom):
    glist = [g.contiguous() for g in geom]
    A, U, V = p.measured.shape
    proj = _native.fp_forward(p.volume, glist, p.pixel_size, U, V, p.step_size, p.voxel_size, 0., 0., 0.)
    r = proj - p.measured
    loss = (r.double() ** 2).mean().item()
    up = (2.0 / (A * U * V)) * r
    _, gg = _native.fp_backward(up.contiguous(), p.volume, glist, p.pixel_size, p.step_size, p.voxel_size,
                                False, 0., 0., 0.)
    return loss, gg


def _rel(a, b):
    return ((a.double() - b.double()).norm() / b.double().norm().clamp_min(1e-30)).item()


@pytest.mark.timeout(900)
def test_batched_calibration_workload_64_problems():
    specs = phantom.calibration_specs(64, seed=4)
    print(f"setting up {len(specs)} problems", flush=True)
    probs = [calibration.synthetic_problem(s, DEV) for s in specs]
    starts = [p.geometry.clone() for p in probs]
    torch.cuda.synchronize()

    # step 0 in one batched graph: loss and gradient against the plain operators
    calibration.calibrate(probs, steps=0, use_graph=True, want_grad=True)
    e_loss, e_grad = [], []
    for p, g0 in zip(probs, starts):
        loss, grad = _plain_loss_grad(p, g0)
        e_loss.append(abs(p.loss_history[0].item() - loss) / loss)
        e_grad.append([_rel(p.grad[k], grad[k]) for k in range(8)])
    e_grad = np.array(e_grad)
    print(f"step 0, fused vs plain: loss max {max(e_loss):.2e}; gradient per group max "
          f"{np.array2string(e_grad.max(axis=0), precision=1)}, median {np.median(e_grad):.1e}", flush=True)
    assert max(e_loss) <= 1e-5
    assert np.median(e_grad) <= 1e-5 and e_grad.max() <= 1e-4

    # the 20-step fit of all 64 problems in one graph
    for p, g0 in zip(probs, starts):
        p.geometry.copy_(g0)
        p.adam_m.zero_()
        p.adam_v.zero_()
    calibration.calibrate(probs, steps=STEPS, use_graph=True)
    size = [s["D"] ** 3 * s["A"] for s in specs]
    order = np.argsort(size)
    picks = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    for i in picks:
        p = probs[i]
        q = calibration.CalibrationProblem(p.volume, p.measured, starts[i], p.pixel_size, p.step_size,
                                           p.voxel_size)
        calibration.calibrate([q], steps=STEPS, use_graph=True)
        print(f"problem {i} ({specs[i]['D']}^3 x {specs[i]['A']} views, {specs[i]['U']}^2): alone == batched "
              f"{torch.equal(q.geometry, p.geometry)}", flush=True)
        assert torch.equal(q.geometry, p.geometry)
        assert torch.allclose(q.loss_history, p.loss_history, rtol=1e-6, atol=0)

    ratio = [p.loss_history[-1].item() / p.loss_history[0].item() for p in probs]
    err0 = [calibration.reprojection_error_px(g0, p.true_geometry, tuple(p.volume.shape), p.voxel_size,
                                              p.pixel_size) for g0, p in zip(starts, probs)]
    err1 = [calibration.reprojection_error_px(p.geometry, p.true_geometry, tuple(p.volume.shape), p.voxel_size,
                                              p.pixel_size) for p in probs]
    print(f"loss ratio median {np.median(ratio):.3f} max {max(ratio):.3f}; reprojection error median "
          f"{np.median(err0):.3f} -> {np.median(err1):.3f} px", flush=True)
    assert max(ratio) < 1.0
    assert np.median(err1) < 0.5 * np.median(err0)

"""Autotuner (GPU): every candidate brick shape passes the parity checks of the
default, and the tuner picks a candidate, caches it and reuses it."""

import json

import numpy as np
import pytest
import torch

import cases
import oracle
from cbct_backprojector import _native, autotune, calibration

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _case():
    D = H = W = 40
    A, U, V = 10, 66, 54
    rng = np.random.default_rng(31)
    vol = torch.tensor(rng.random((D, H, W), dtype=np.float32), device=DEV)
    geom = [torch.tensor(x, device=DEV) for x in (
        np.linspace(0.2, 6.1, A).astype(np.float32), np.full(A, 230.0, np.float32),
        np.full(A, 130.0, np.float32), np.full(A, 32.5, np.float32) + rng.uniform(-2, 2, A).astype(np.float32),
        np.full(A, 26.5, np.float32), rng.uniform(-1, 1, A).astype(np.float32),
        rng.uniform(-1, 1, A).astype(np.float32), rng.uniform(-1, 1, A).astype(np.float32))]
    up = torch.tensor(rng.standard_normal((A, U, V)).astype(np.float32), device=DEV)
    return vol, geom, up, (0.8, U, V, 0.25, 0.5)


@pytest.fixture(scope="module")
def reference():
    vol, geom, up, (pix, U, V, step, vs) = _case()
    ref32 = oracle.fp_forward(cases.to_np(vol), [cases.to_np(g) for g in geom], pix, U, V, step, vs, mode="f32")
    with _native.fp_tuning(1, *autotune.DEFAULT):
        gv, gg = _native.fp_backward(up, vol, geom, pix, step, vs, True, 0., 0., 0.)
    torch.cuda.synchronize()
    return ref32, cases.to_np(gv), cases.to_np(gg)


@pytest.mark.parametrize("tuning", autotune.candidates())
def test_every_candidate_passes_projector_parity(tuning, reference):
    ref32, ref_gv, ref_gg = reference
    vol, geom, up, (pix, U, V, step, vs) = _case()
    with _native.fp_tuning(1, *tuning):
        p = _native.fp_forward(vol, geom, pix, U, V, step, vs, 0., 0., 0.)
        gv, gg = _native.fp_backward(up, vol, geom, pix, step, vs, True, 0., 0., 0.)
        _, gg0 = _native.fp_backward(up, vol, geom, pix, step, vs, False, 0., 0., 0.)
    torch.cuda.synchronize()
    assert cases.rel_l2(cases.to_np(p), ref32) <= 1e-6
    assert torch.equal(gg0, gg)
    assert cases.rel_l2(cases.to_np(gg), ref_gg) <= 1e-6
    assert cases.rel_l2(cases.to_np(gv), ref_gv) <= 2e-6


@pytest.mark.parametrize("tuning", autotune.candidates())
def test_every_candidate_gives_the_same_calibration_gradient(tuning):
    spec = {"D": 36, "A": 8, "U": 48, "V": 40, "voxel": 0.5, "sod": 600.0, "sid": 1100.0, "seed": 3,
            "pixel": 0.75}
    ref = calibration.synthetic_problem(spec, DEV, tuning=autotune.DEFAULT)
    calibration.calibrate([ref], steps=0, use_graph=False, want_grad=True)
    p = calibration.synthetic_problem(spec, DEV, tuning=tuning)
    calibration.calibrate([p], steps=0, use_graph=False, want_grad=True)
    assert abs(p.loss_history[0].item() - ref.loss_history[0].item()) <= 1e-6 * ref.loss_history[0].item()
    for k in range(8):
        assert cases.rel_l2(cases.to_np(p.grad[k]), cases.to_np(ref.grad[k])) <= 1e-6


def test_tuner_picks_caches_and_reuses(tmp_path, monkeypatch):
    monkeypatch.setenv("CBCT_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    vol, geom, _, (pix, U, V, step, vs) = _case()
    best = autotune.tune_projector(vol, geom, pix, U, V, step, vs, op="adjoint", views=4, reps=1)
    assert best in autotune.candidates()
    data = json.loads((tmp_path / "tune.json").read_text())
    key = autotune.shape_key("adjoint", vol.shape, U, V, step, vs, 230.0, 130.0)
    assert data[key]["best"] == list(best) and data[key]["ms_per_view"] > 0
    # the cached shape class returns its stored best
    assert autotune.tune_projector(vol, geom, pix, U, V, step, vs, op="adjoint") == best


def test_calibration_tuner_sets_problem_tuning(tmp_path, monkeypatch):
    monkeypatch.setenv("CBCT_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    spec = {"D": 36, "A": 12, "U": 48, "V": 48, "voxel": 0.5, "sod": 600.0, "sid": 1100.0, "seed": 4,
            "pixel": 0.75}
    p = calibration.synthetic_problem(spec, DEV)
    best = autotune.tune_calibration(p, views=4, reps=1)
    assert p.tuning == best and best in autotune.batch_candidates()
    calibration.calibrate([p], steps=2)
    assert p.loss_history[2].item() < p.loss_history[0].item()

"""Per-call options and the geometry-off adjoint (GPU).

* cbct_fp_backward_ex without grad_geometry gives the volume gradient of the full adjoint:
  bitwise in deterministic mode (int64 accumulation is order-free), to float-atomic rounding
  otherwise, and within the FP32 oracle's tolerance; it runs for every kernel setting.
* The autograd node asks for the geometry gradients only when a geometry tensor needs one,
  and for the volume gradient only when the volume needs one; both paths match the oracle.
* Settings captured when an operator's forward ran govern its backward, even when the backward
  is run after the thread's settings changed.
"""

import numpy as np
import pytest
import torch

import cases
import oracle
from cbct_backprojector import _native, project

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _case(seed=43):
    rng = np.random.default_rng(seed)
    D, H, W = 36, 40, 44
    A, U, V = 12, 64, 48
    vol = rng.random((D, H, W), dtype=np.float32)
    geom = [np.linspace(0.2, 6.1, A).astype(np.float32), np.full(A, 210.0, np.float32),
            np.full(A, 130.0, np.float32), (31.5 + rng.uniform(-2, 2, A)).astype(np.float32),
            np.full(A, 23.5, np.float32), rng.uniform(-1, 1, A).astype(np.float32),
            rng.uniform(-1, 1, A).astype(np.float32), rng.uniform(-1, 1, A).astype(np.float32)]
    up = rng.standard_normal((A, U, V)).astype(np.float32)
    return vol, geom, up, (0.9, 0.3, 0.5), (0.2, -0.1, 0.15)


def _dev(vol, geom, up):
    return torch.tensor(vol, device=DEV), [torch.tensor(g, device=DEV) for g in geom], torch.tensor(up, device=DEV)


@pytest.mark.parametrize("tuning", [None, (1, 12, 3840, 1), (1, 8, 4096, 0), (1, 4, 2048, 1), (0, 12, 3840, 1)])
def test_geometry_off_adjoint_matches_full_adjoint(tuning):
    vol, geom, up, (pix, step, vs), off = _case()
    tv, tg, tu = _dev(vol, geom, up)
    print(f"geometry-off adjoint, tuning {tuning}", flush=True)
    opts = _native.options(_native.TUNING_SCATTER, tuning=tuning, deterministic=False)
    gv_full, gg_full = _native.fp_backward(tu, tv, tg, pix, step, vs, True, *off, opts=opts)
    gv_vol, gg_none = _native.fp_backward(tu, tv, tg, pix, step, vs, True, *off, compute_grad_geometry=False,
                                          opts=opts)
    assert gg_none is None
    assert cases.rel_l2(cases.to_np(gv_vol), cases.to_np(gv_full)) <= 1e-6
    rvol, _ = oracle.fp_backward(up, vol, geom, pix, step, vs, offset=off, need_volume=True, mode="f32")
    assert cases.rel_l2(cases.to_np(gv_vol), rvol) <= 1e-5
    if tuning is None or tuning[0] == 1:
        det = _native.options(_native.TUNING_SCATTER, tuning=tuning, deterministic=True)
        a, _ = _native.fp_backward(tu, tv, tg, pix, step, vs, True, *off, opts=det)
        b, _ = _native.fp_backward(tu, tv, tg, pix, step, vs, True, *off, compute_grad_geometry=False, opts=det)
        assert torch.equal(a, b)


def test_autograd_requests_only_the_needed_gradients():
    vol, geom, up, (pix, step, vs), off = _case(44)
    tv, tg, tu = _dev(vol, geom, up)
    # the FP32 schedule: at step 0.3 (not a power of two) the FP64 schedule's sample positions drift
    # away from the accumulated FP32 ones, which moves the scattered weights
    rvol, rgeo = oracle.fp_backward(up, vol, geom, pix, step, vs, offset=off, mode="f32")
    # volume only: the geometry-off kernel
    v = tv.clone().requires_grad_(True)
    proj = project(v, *tg, pix, 64, 48, step, vs, *off)
    (gv,) = torch.autograd.grad(proj, [v], grad_outputs=tu)
    assert cases.rel_l2(cases.to_np(gv), rvol) <= 1e-5
    # geometry only: the gathering adjoint
    g = [x.clone().requires_grad_(True) for x in tg]
    proj = project(tv, *g, pix, 64, 48, step, vs, *off)
    grads = torch.autograd.grad(proj, g, grad_outputs=tu)
    for k in range(8):
        ref = rgeo[k]
        got = cases.to_np(grads[k])
        assert np.allclose(got, ref, rtol=2e-3, atol=2e-3 * max(1.0, float(np.abs(ref).max()))), k


def test_settings_captured_at_forward_govern_backward():
    """A projection made in deterministic mode gives a bitwise reproducible volume gradient even
    when backward() runs after the thread left deterministic mode."""
    vol, geom, up, (pix, step, vs), off = _case(45)
    tv, tg, tu = _dev(vol, geom, up)
    grads = []
    for _ in range(2):
        v = tv.clone().requires_grad_(True)
        with _native.deterministic():
            proj = project(v, *tg, pix, 64, 48, step, vs, *off)
        assert not _native.get_deterministic()
        (g,) = torch.autograd.grad(proj, [v], grad_outputs=tu)
        grads.append(g)
    assert torch.equal(grads[0], grads[1])
    with _native.deterministic():
        ref, _ = _native.fp_backward(tu, tv, tg, pix, step, vs, True, *off)
    assert torch.equal(grads[0], ref)

"""Deterministic mode of the projector adjoint (GPU): the volume gradient accumulated as
int64 fixed point is bitwise identical run to run, agrees with the float-atomic result and
the FP32 oracle, and leaves the geometry gradients unchanged."""

import numpy as np
import pytest
import torch

import bench
import cases
import oracle
from cbct_backprojector import _native

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _case():
    rng = np.random.default_rng(41)
    D = H = W = 40
    A, U, V = 16, 70, 52
    vol = rng.random((D, H, W), dtype=np.float32)
    geom = [np.linspace(0.1, 6.0, A).astype(np.float32), np.full(A, 220.0, np.float32),
            np.full(A, 140.0, np.float32), (34.5 + rng.uniform(-2, 2, A)).astype(np.float32),
            np.full(A, 25.5, np.float32), rng.uniform(-1, 1, A).astype(np.float32),
            rng.uniform(-1, 1, A).astype(np.float32), rng.uniform(-1, 1, A).astype(np.float32)]
    up = rng.standard_normal((A, U, V)).astype(np.float32)
    return vol, geom, up, (0.9, 0.25, 0.5)


def test_deterministic_volume_gradient_is_reproducible_and_matches():
    vol, geom, up, (pix, step, vs) = _case()
    tv, tg, tu = torch.tensor(vol, device=DEV), [torch.tensor(g, device=DEV) for g in geom], torch.tensor(up, device=DEV)
    args = (pix, step, vs, True, 0.3, -0.2, 0.1)
    gv_f, gg_f = _native.fp_backward(tu, tv, tg, *args)
    with _native.deterministic():
        runs = [_native.fp_backward(tu, tv, tg, *args) for _ in range(3)]
    for gv, gg in runs[1:]:
        assert torch.equal(gv, runs[0][0])
        assert torch.equal(gg, runs[0][1])
    gv_d, gg_d = runs[0]
    assert torch.equal(gg_d, gg_f)
    assert cases.rel_l2(cases.to_np(gv_d), cases.to_np(gv_f)) <= 1e-6
    rvol, _ = oracle.fp_backward(up, vol, geom, pix, step, vs, offset=(0.3, -0.2, 0.1), mode="f32")
    err = cases.rel_l2(cases.to_np(gv_d), rvol)
    print(f"deterministic volume gradient vs FP32 oracle: {err:.2e}")
    assert err <= 1e-5
    assert not _native.get_deterministic()


def test_deterministic_mode_at_full_size_subset():
    w = bench.WORKLOADS["geometry-gradients-256c-720v"]
    vol_np, geom_np, vol, geom, _ = bench.make_inputs(w, DEV)
    views = torch.tensor([3, 300, 611], device=DEV)
    g = [x[views].contiguous() for x in geom]
    gen = torch.Generator(device=DEV).manual_seed(5)
    up = torch.randn((3, w["U"], w["V"]), generator=gen, device=DEV)
    args = (w["pixel"], w["step"], w["voxel"], True, 0., 0., 0.)
    gv_f, _ = _native.fp_backward(up, vol, g, *args)
    with _native.deterministic():
        a, _ = _native.fp_backward(up, vol, g, *args)
        b, _ = _native.fp_backward(up, vol, g, *args)
    assert torch.equal(a, b)
    assert cases.rel_l2(cases.to_np(a), cases.to_np(gv_f)) <= 1e-6


def test_deterministic_mode_needs_the_slab_kernels():
    vol, geom, up, (pix, step, vs) = _case()
    tv, tg, tu = torch.tensor(vol, device=DEV), [torch.tensor(g, device=DEV) for g in geom], torch.tensor(up, device=DEV)
    with _native.deterministic(), _native.fp_tuning(0):
        with pytest.raises(RuntimeError, match="slab kernels"):
            _native.fp_backward(tu, tv, tg, pix, step, vs, True, 0., 0., 0.)

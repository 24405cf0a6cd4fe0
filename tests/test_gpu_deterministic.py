"""Deterministic mode (GPU).  Projector adjoint: the volume gradient accumulated as int64
fixed point is bitwise identical run to run, agrees with the float-atomic result and the FP32
oracle, and leaves the geometry gradients unchanged.  Backprojection backward: the geometry
gradients are reproducible in both modes (fixed-order reduction); in deterministic mode the
projection gradient is too, on the LDS-tile path and on the direct global path."""

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


def _bp_case(pixel):
    rng = np.random.default_rng(23)
    D = H = W = 24
    A = 5
    U = V = 40 if pixel > 0.5 else 700
    geom = [np.linspace(0.2, 5.9, A).astype(np.float32), np.full(A, 260.0, np.float32),
            np.full(A, 150.0, np.float32), ((U - 1) / 2 + rng.uniform(-2, 2, A)).astype(np.float32),
            ((V - 1) / 2 + rng.uniform(-2, 2, A)).astype(np.float32), rng.uniform(-1, 1, A).astype(np.float32),
            rng.uniform(-1, 1, A).astype(np.float32), rng.uniform(-1, 1, A).astype(np.float32)]
    proj = rng.random((A, U, V)).astype(np.float32)
    gv = rng.standard_normal((D, H, W)).astype(np.float32)
    return gv, proj, geom


@pytest.mark.parametrize("pixel", [0.9, 0.04])
def test_deterministic_bp_backward_is_reproducible_and_matches(pixel):
    """pixel 0.9: LDS-tile path; pixel 0.04: windows larger than the tile, direct int64 atomics."""
    gv, proj, geom = _bp_case(pixel)
    tgv, tp = torch.tensor(gv, device=DEV), torch.tensor(proj, device=DEV)
    tg = [torch.tensor(g, device=DEV) for g in geom]
    args = (pixel, 0.7, 0.1, -0.2, 0.3)
    float_runs = [_native.bp_backward(tgv, tp, tg, *args) for _ in range(2)]
    assert torch.equal(float_runs[0][1], float_runs[1][1])        # geometry: fixed-order sums
    with _native.deterministic():
        runs = [_native.bp_backward(tgv, tp, tg, *args) for _ in range(3)]
    for gp, gg in runs[1:]:
        assert torch.equal(gp, runs[0][0])
        assert torch.equal(gg, runs[0][1])
    gp_d, gg_d = runs[0]
    assert torch.equal(gg_d, float_runs[0][1])
    e_float = cases.rel_l2(cases.to_np(gp_d), cases.to_np(float_runs[0][0]))
    rproj, _ = oracle.bp_backward(gv.astype(np.float64), proj.astype(np.float64), geom, pixel, 0.7,
                                  offset=(0.1, -0.2, 0.3))
    e_oracle = cases.rel_l2(cases.to_np(gp_d), rproj)
    print(f"deterministic bp backward pixel {pixel}: vs float path {e_float:.2e}, vs FP64 oracle {e_oracle:.2e}")
    assert e_float <= 1e-6 and e_oracle <= 1e-5
    assert not _native.get_deterministic()


def test_deterministic_bp_backward_at_full_size_subset():
    from cbct_backprojector import phantom
    w = bench.WORKLOADS[bench.BP_WORKLOAD]
    gen = torch.Generator(device=DEV).manual_seed(9)
    D = w["D"]
    views = [0, 97, 250]
    geom = phantom.perturb_geometry(phantom.circular_geometry(w["A"], w["sid"], w["sod"], w["U"], w["V"],
                                                              w["pixel"]), bench.PERTURB_SEED)
    g = [geom[k][views].contiguous().to(DEV) for k in phantom.GEOMETRY_NAMES]
    proj = torch.rand((len(views), w["U"], w["V"]), generator=gen, device=DEV)
    gvol = torch.randn((D, D, D), generator=gen, device=DEV)
    args = (w["pixel"], w["voxel"], 0.0, 0.0, 0.0)
    gp_f, gg_f = _native.bp_backward(gvol, proj, g, *args)
    with _native.deterministic():
        a = _native.bp_backward(gvol, proj, g, *args)
        b = _native.bp_backward(gvol, proj, g, *args)
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
    assert torch.equal(a[1], gg_f)
    assert cases.rel_l2(cases.to_np(a[0]), cases.to_np(gp_f)) <= 1e-6

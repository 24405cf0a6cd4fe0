"""Edge cases of the projector and backprojector (GPU) against the CPU oracle.

* Rays with exactly zero direction components (odd detectors whose centre pixel sits on the
  central ray, views at beta = 0 and pi, no source offsets): the slab test's "parallel" branch
  (csrc/src_cbct_forwardprojector_cuda.cu:55-67) and samples on voxel boundaries.
* Detectors of one row or one column (project() accepts U, V >= 1, source_only_forward_projector.py:53-57)
  and a single view.
* Views whose rays mostly or entirely miss the volume.
The default staged kernels must agree bit for bit with the direct kernel on the projections.
"""

import math

import numpy as np
import pytest
import torch

import cases
import oracle
from cbct_backprojector import _native

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _geometry(angles, U, V, sid=150.0, sod=90.0, du=0.0, dv=0.0):
    A = len(angles)
    return [np.asarray(angles, np.float32), np.full(A, sid, np.float32), np.full(A, sod, np.float32),
            np.full(A, (U - 1) / 2 + du, np.float32), np.full(A, (V - 1) / 2 + dv, np.float32),
            np.zeros(A, np.float32), np.zeros(A, np.float32), np.zeros(A, np.float32)]


def _run_fp(vol, geom, pix, U, V, step, vs, off, up):
    tv = torch.tensor(vol, device=DEV)
    tg = [torch.tensor(g, device=DEV) for g in geom]
    p = _native.fp_forward(tv, tg, pix, U, V, step, vs, *off)
    gv, gg = _native.fp_backward(torch.tensor(up, device=DEV), tv, tg, pix, step, vs, True, *off)
    torch.cuda.synchronize()
    return cases.to_np(p), cases.to_np(gv), cases.to_np(gg)


def _check_fp(vol, geom, pix, U, V, step, vs, off, seed):
    rng = np.random.default_rng(seed)
    up = rng.standard_normal((len(geom[0]), U, V)).astype(np.float32)
    p, gv, gg = _run_fp(vol, geom, pix, U, V, step, vs, off, up)
    with _native.fp_tuning(0):
        p_direct, _, _ = _run_fp(vol, geom, pix, U, V, step, vs, off, up)
    assert np.array_equal(p, p_direct)
    ref32 = oracle.fp_forward(vol, geom, pix, U, V, step, vs, off, mode="f32")
    ref64 = oracle.fp_forward(vol.astype(np.float64), geom, pix, U, V, step, vs, off, mode="f64")
    scale = max(np.abs(ref64).max(), 1e-30)
    assert np.abs(p - ref32).max() <= 1e-6 * scale
    np.testing.assert_allclose(p, ref64, rtol=5e-5, atol=5e-5 * scale)
    rvol, rgeo = oracle.fp_backward(up, vol, geom, pix, step, vs, off, mode="f32")
    _, rgeo64 = oracle.fp_backward(up.astype(np.float64), vol.astype(np.float64), geom, pix, step, vs, off,
                                   mode="f64")
    assert cases.rel_l2(gv, rvol) <= 1e-5
    # A geometry gradient is a sum with cancellation over few rays here, so its FP32 error is
    # compared with the FP64 oracle: within 1e-4 of its size (the reference's own bar is 1e-3,
    # examples/check_forward_operator.py:226) or at most twice the FP32 oracle's error.
    for k in range(8):
        e_gpu = np.abs(gg[k] - rgeo64[k]).max()
        e_f32 = np.abs(rgeo[k] - rgeo64[k]).max()
        assert e_gpu <= max(1e-4 * max(np.abs(rgeo64[k]).max(), 1e-30), 2.0 * e_f32), k
    return p, gv, gg


def test_axis_parallel_rays_and_voxel_boundary_samples():
    """Odd detector with integer centres: the centre row has d_z == 0 and, at beta = 0 and pi, the
    centre column d_x == 0 (sin 0 == 0); the even volume puts those rays on voxel faces."""
    rng = np.random.default_rng(3)
    D, H, W = 20, 24, 22
    U, V = 33, 31
    vol = rng.random((D, H, W), dtype=np.float32)
    geom = _geometry([0.0, math.pi / 2, math.pi, 1.0, 4.0], U, V)
    p, _, _ = _check_fp(vol, geom, 1.0, U, V, 0.5, 1.0, (0.0, 0.0, 0.0), seed=4)
    ray = oracle.fp_forward(vol.astype(np.float64), geom, 1.0, U, V, 0.5, 1.0, (0.0, 0.0, 0.0), mode="f64")
    assert np.count_nonzero(ray[:, :, (V - 1) // 2]) > 0          # the z-parallel row hits the volume
    assert np.count_nonzero(p[0, (U - 1) // 2, :]) > 0              # and so does the x-parallel column


def test_one_row_one_column_and_one_view_detectors():
    rng = np.random.default_rng(5)
    vol = rng.random((16, 20, 24), dtype=np.float32)
    for U, V in ((1, 40), (48, 1), (1, 1)):
        geom = _geometry([0.7], U, V, sid=180.0, sod=120.0, du=0.3, dv=-0.2)
        _check_fp(vol, geom, 1.0, U, V, 0.5, 1.0, (0.1, -0.2, 0.3), seed=U * 100 + V)


def test_views_that_miss_the_volume():
    """Detector centres shifted far off: every ray of view 1 misses (zero projection, zero
    gradients), view 0 is only partly hit."""
    rng = np.random.default_rng(6)
    vol = rng.random((16, 16, 16), dtype=np.float32)
    U, V = 40, 36
    geom = _geometry([0.3, 2.0], U, V)
    geom[3] = np.array([(U - 1) / 2 + 25.0, (U - 1) / 2 + 400.0], np.float32)
    p, gv, gg = _check_fp(vol, geom, 1.0, U, V, 0.5, 1.0, (0.0, 0.0, 0.0), seed=7)
    assert not p[1].any() and not gg[:, 1].any()
    assert p[0].any() and not p[0].all()


def test_default_staging_steep_cone_matches_direct_kernel():
    """Library default tunings on a steep cone (source 35 mm from the axis of a 20x22x18 mm volume,
    ~29 degree half-angle) with source offsets: oblique rays in every axis and slab boxes that hit
    the LDS capacity, so lateral box edges, halving and the per-lane in-box proof (lanes that fail it
    take the global path) all run.  Projections equal the direct kernel's bit for bit."""
    rng = np.random.default_rng(21)
    D, H, W = 36, 44, 40
    A, U, V = 8, 64, 64
    vol = torch.tensor(rng.random((D, H, W), dtype=np.float32), device=DEV)
    geom_np = [np.linspace(0.2, 5.9, A).astype(np.float32), np.full(A, 70.0, np.float32),
               np.full(A, 35.0, np.float32), (31.5 + rng.uniform(-3, 3, A)).astype(np.float32),
               (31.5 + rng.uniform(-3, 3, A)).astype(np.float32), rng.uniform(-1, 1, A).astype(np.float32),
               rng.uniform(-1, 1, A).astype(np.float32), rng.uniform(-1, 1, A).astype(np.float32)]
    geom = [torch.tensor(g, device=DEV) for g in geom_np]
    up = torch.tensor(rng.standard_normal((A, U, V)).astype(np.float32), device=DEV)
    pix, step, vs, off = 1.2, 0.25, 0.5, (0.4, -0.3, 0.2)
    p = _native.fp_forward(vol, geom, pix, U, V, step, vs, *off)
    gv, gg = _native.fp_backward(up, vol, geom, pix, step, vs, True, *off)
    with _native.fp_tuning(0):
        p_d = _native.fp_forward(vol, geom, pix, U, V, step, vs, *off)
        gv_d, gg_d = _native.fp_backward(up, vol, geom, pix, step, vs, True, *off)
    torch.cuda.synchronize()
    assert torch.equal(p, p_d)
    assert cases.rel_l2(cases.to_np(gg), cases.to_np(gg_d)) <= 1e-6
    assert cases.rel_l2(cases.to_np(gv), cases.to_np(gv_d)) <= 2e-6
    ref32 = oracle.fp_forward(cases.to_np(vol), geom_np, pix, U, V, step, vs, off, mode="f32")
    assert cases.rel_l2(cases.to_np(p), ref32) <= 1e-6
    assert np.count_nonzero(ref32) > 0.4 * ref32.size      # about half the rays cross the volume


def test_backprojection_odd_detector_and_offset_volume():
    rng = np.random.default_rng(8)
    D, H, W = 9, 14, 11
    U, V = 33, 31
    A = 4
    geom = _geometry([0.0, math.pi / 2, 2.5, math.pi], U, V, sid=160.0, sod=100.0)
    geom[5] = rng.uniform(-1, 1, A).astype(np.float32)
    proj = rng.random((A, U, V)).astype(np.float32)
    gvol = rng.standard_normal((D, H, W)).astype(np.float32)
    off = (0.4, -0.3, 0.2)
    tg = [torch.tensor(g, device=DEV) for g in geom]
    vol = _native.bp_forward(torch.tensor(proj, device=DEV), tg, 0.9, D, H, W, 0.7, *off)
    gp, gg = _native.bp_backward(torch.tensor(gvol, device=DEV), torch.tensor(proj, device=DEV), tg, 0.9, 0.7, *off)
    torch.cuda.synchronize()
    rvol = oracle.bp_forward(proj, geom, 0.9, (D, H, W), 0.7, offset=off, mode="f32")
    assert cases.rel_l2(cases.to_np(vol), rvol) <= 1e-6
    rgp, rgg = oracle.bp_backward(gvol.astype(np.float64), proj.astype(np.float64), geom, 0.9, 0.7, offset=off,
                                  mode="f64")
    assert cases.rel_l2(cases.to_np(gp), rgp) <= 1e-5
    for k in range(8):
        assert cases.rel_l2(cases.to_np(gg[k]), rgg[k]) <= 1e-5, k

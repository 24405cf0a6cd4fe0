"""A volume with more than 2^31 voxels (GPU): the forward projector and its adjoint index the
volume with 64-bit offsets, as the reference's forward kernel does
(csrc/src_cbct_forwardprojector_cuda.cu:200-208).

The volume is 1296^3 fp32 (2.18e9 voxels, 8.7 GB); only its last 26 z-planes are non-zero, and
the volume is shifted along z so that the rays of a small detector cross exactly those planes,
whose voxel offsets all lie beyond 2^31.  Projections, the volume gradient (of the non-zero
planes, the only ones rays reach) and the geometry gradients are compared with the FP32
oracle, which indexes in 64 bits; the slab kernels and the direct kernels are both run."""

import numpy as np
import pytest
import torch

import cases
import oracle
from cbct_backprojector import _native

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
N = 1296
Z0 = 1270          # first non-zero plane; plane 1279 starts beyond 2^31


@pytest.mark.timeout(900)
def test_projector_on_a_volume_beyond_2_31_voxels():
    assert N ** 3 > 2 ** 31 and 1279 * N * N > 2 ** 31
    rng = np.random.default_rng(77)
    vol_np = np.zeros((N, N, N), np.float32)
    vol_np[Z0:] = rng.random((N - Z0, N, N), dtype=np.float32)
    vs, step, pix = 0.5, 0.25, 0.5
    off = (0.0, 0.0, -(1285 - N / 2) * vs)        # voxel plane 1285 at world z = 0
    A, U, V = 2, 8, 8
    geom = [np.array([0.3, 1.9], np.float32), np.full(A, 3000.0, np.float32), np.full(A, 2000.0, np.float32),
            np.full(A, 3.5, np.float32), np.full(A, 3.5, np.float32), np.array([0.2, -0.1], np.float32),
            np.array([0.1, 0.3], np.float32), np.array([-0.2, 0.15], np.float32)]
    up = rng.standard_normal((A, U, V)).astype(np.float32)
    print("uploading a 1296^3 volume", flush=True)
    vol = torch.from_numpy(vol_np).to(DEV)
    g = [torch.tensor(x, device=DEV) for x in geom]
    tu = torch.tensor(up, device=DEV)
    ref = oracle.fp_forward(vol_np, geom, pix, U, V, step, vs, offset=off, mode="f32", threads=1)
    assert np.abs(ref).min() > 0          # every ray crosses the non-zero planes
    rvol, rgeo = oracle.fp_backward(up, vol_np, geom, pix, step, vs, offset=off, mode="f32", threads=1)
    for tuning in ((1, 12, 5120, 0), (0, 12, 5120, 0)):
        print(f"tuning {tuning}", flush=True)
        fo = _native.options(_native.TUNING_GATHER, tuning=tuning, deterministic=False)
        proj = _native.fp_forward(vol, g, pix, U, V, step, vs, *off, opts=fo)
        e = cases.rel_l2(cases.to_np(proj), ref)
        bo = _native.options(_native.TUNING_SCATTER, tuning=tuning[:3] + (1,), deterministic=False)
        gv, gg = _native.fp_backward(tu, vol, g, pix, step, vs, True, *off, opts=bo)
        slab = gv[Z0:].cpu().numpy()
        ev = cases.rel_l2(slab, rvol[Z0:])
        outside = float(gv[:Z0 - 2].abs().max())
        print(f"  forward rel L2 {e:.2e}; volume gradient rel L2 {ev:.2e}; max |gradient| below the planes "
              f"{outside:.1e}", flush=True)
        assert e <= 1e-6
        assert ev <= 1e-5 and outside == 0.0
        # geometry only (the gathering adjoint) and volume only (no geometry sums)
        _, gg1 = _native.fp_backward(tu, vol, g, pix, step, vs, False, *off, opts=fo)
        gv2, _ = _native.fp_backward(tu, vol, g, pix, step, vs, True, *off, compute_grad_geometry=False, opts=bo)
        assert cases.rel_l2(gv2[Z0:].cpu().numpy(), rvol[Z0:]) <= 1e-5
        for k in range(8):
            r = rgeo[k]
            tol = dict(rtol=1e-3, atol=1e-3 * max(1.0, float(np.abs(r).max())))
            assert np.allclose(cases.to_np(gg[k]), r, **tol), k
            assert np.allclose(cases.to_np(gg1[k]), r, **tol), k
        del gv, gv2
    del vol
    torch.cuda.empty_cache()

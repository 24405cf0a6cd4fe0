"""Thin volumes with wide planes (GPU).  The slab kernels form the staging and flush offsets of a
box that lies inside the volume by 24-bit multiply-adds, which is exact only while 4 H W < 2^24;
wider planes take the clamped decomposition for every box.  Volumes of 64 planes on either side of
that limit (2040^2 = 4,161,600 and 2100^2 = 4,410,000 voxels per plane, against 2^22 = 4,194,304)
give projections bit-identical to the direct kernel (the same per-sample arithmetic), the same
geometry gradients up to the order of the per-tile sums, and the same volume gradient up to
atomic ordering and the adjoint's fixed-point box.

The rays cross the planes' full width (2040-2100 voxels) with the detector's central 256 x 64
pixels, so most slab boxes lie inside the volume and the first, last and top and bottom ones
touch its faces."""

import numpy as np
import pytest
import torch

import cases
from cbct_backprojector import _native

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


@pytest.mark.parametrize("side", [2040, 2100])
def test_wide_plane_volumes_match_the_direct_kernel(side):
    D = 64
    assert (side * side < 2 ** 22) == (side == 2040)
    gen = torch.Generator(device=DEV).manual_seed(side)
    vol = torch.rand((D, side, side), generator=gen, device=DEV)
    vs, step, pix = 0.5, 0.25, 1.0
    A, U, V = 3, 256, 64
    geom = [torch.tensor(x, device=DEV) for x in (
        np.array([0.2, 1.1, 2.5], np.float32), np.full(A, 3000.0, np.float32), np.full(A, 1500.0, np.float32),
        np.full(A, 127.5, np.float32), np.full(A, 31.5, np.float32), np.array([0.3, -0.2, 0.1], np.float32),
        np.array([-0.1, 0.2, 0.3], np.float32), np.array([0.05, -0.1, 0.2], np.float32))]
    up = torch.randn((A, U, V), generator=gen, device=DEV)
    off = (0.0, 0.0, 0.0)
    out = {}
    for name, use_slab in (("direct", 0), ("slab", 1)):
        print(f"side {side}: {name} kernels", flush=True)
        fo = _native.options(_native.TUNING_GATHER, tuning=(use_slab, 12, 5120, 0), deterministic=False)
        bo = _native.options(_native.TUNING_SCATTER, tuning=(use_slab, 12, 3840, 1), deterministic=False)
        proj = _native.fp_forward(vol, geom, pix, U, V, step, vs, *off, opts=fo)
        gv, gg = _native.fp_backward(up, vol, geom, pix, step, vs, True, *off, opts=bo)
        _, gg1 = _native.fp_backward(up, vol, geom, pix, step, vs, False, *off, opts=fo)
        out[name] = (proj, gv, gg, gg1)
    torch.cuda.synchronize()
    p0, gv0, gg0, g10 = out["direct"]
    p1, gv1, gg1, g11 = out["slab"]
    hit = float((p0 > 0).float().mean())
    ev = float((gv1 - gv0).norm() / gv0.norm())
    eg = max(cases.rel_l2(cases.to_np(gg1[k]), cases.to_np(gg0[k])) for k in range(8))
    print(f"side {side}: rays that meet the volume {hit:.3f}; volume gradient rel L2 {ev:.2e}; "
          f"geometry gradients rel L2 {eg:.2e}", flush=True)
    assert hit > 0.9
    assert torch.equal(p1, p0)
    # the gathering adjoint uses the forward's 8x8 lane blocks: the same sums in the same order
    assert torch.equal(g11, g10)
    assert eg <= 1e-6
    # (1.7e-6 measured: the fixed-point box's rounding; a wrong offset would be of order 1)
    assert ev <= 5e-6
    del vol, gv0, gv1
    torch.cuda.empty_cache()

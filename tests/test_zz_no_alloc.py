"""The allocation-free adjoint (cbct_fp_backward_noalloc) against the workspace entry point
cbct_fp_backward_ex (GPU), called through the C ABI: two calls in a row, the second on a smaller
input, each compared with cbct_fp_backward_ex on the same inputs -- geometry gradients bit for
bit (both reduce in a fixed order), the volume gradient bit for bit in deterministic mode and
to float-atomic rounding otherwise.  A size whose workspace exceeds the fixed scratch is refused
before any device work (that check is also run on the CPU)."""

import ctypes

import numpy as np
import pytest
import torch

import cases
from cbct_backprojector import _native

DEV = torch.device("cuda:0")


def _case(seed, D, A, U, V):
    rng = np.random.default_rng(seed)
    vol = rng.random((D, D + 2, D + 4), dtype=np.float32)
    geom = [np.linspace(0.1, 6.0, A).astype(np.float32), np.full(A, 220.0, np.float32),
            np.full(A, 140.0, np.float32), ((U - 1) / 2 + rng.uniform(-2, 2, A)).astype(np.float32),
            np.full(A, (V - 1) / 2, np.float32), rng.uniform(-1, 1, A).astype(np.float32),
            rng.uniform(-1, 1, A).astype(np.float32), rng.uniform(-1, 1, A).astype(np.float32)]
    up = rng.standard_normal((A, U, V)).astype(np.float32)
    return (torch.tensor(vol, device=DEV), [torch.tensor(g, device=DEV) for g in geom],
            torch.tensor(up, device=DEV))


def _compare(vol, geom, up, det):
    args = (0.9, 0.25, 0.5, True, 0.3, -0.2, 0.1)
    opts = _native.options(_native.TUNING_SCATTER, deterministic=det)
    gv_ref, gg_ref = _native.fp_backward(up, vol, geom, *args, opts=opts)
    gv = torch.full_like(gv_ref, float("nan"))
    gg = torch.full_like(gg_ref, float("nan"))
    _native.fp_backward_noalloc(up, vol, geom, *args, grad_volume=gv, grad_geometry=gg, opts=opts)
    torch.cuda.synchronize()
    assert torch.equal(gg, gg_ref)
    if det:
        assert torch.equal(gv, gv_ref)
    else:
        assert cases.rel_l2(cases.to_np(gv), cases.to_np(gv_ref)) <= 1e-6
    # geometry only (the gathering adjoint) and volume only (the geometry-off adjoint)
    g_only = torch.full_like(gg_ref, float("nan"))
    _native.fp_backward_noalloc(up, vol, geom, *args[:3], False, *args[4:], grad_geometry=g_only,
                                opts=_native.options(_native.TUNING_GATHER))
    _, gg_g = _native.fp_backward(up, vol, geom, *args[:3], False, *args[4:])
    assert torch.equal(g_only, gg_g)
    v_only = torch.full_like(gv_ref, float("nan"))
    _native.fp_backward_noalloc(up, vol, geom, *args, grad_volume=v_only, opts=opts)
    torch.cuda.synchronize()
    if det:
        assert torch.equal(v_only, gv_ref)
    else:
        assert cases.rel_l2(cases.to_np(v_only), cases.to_np(gv_ref)) <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("det", [False, True])
def test_noalloc_adjoint_equals_workspace_adjoint_twice(det):
    print(f"no-alloc adjoint, deterministic={det}: large case", flush=True)
    _compare(*_case(71, 48, 40, 96, 80), det)
    print(f"no-alloc adjoint, deterministic={det}: smaller case", flush=True)
    _compare(*_case(72, 20, 9, 33, 27), det)


def test_noalloc_refuses_a_workspace_beyond_the_scratch():
    """2000 views of a 768x640 detector need 123 MB of per-tile partials: refused with
    CBCT_EWORKSPACE before any pointer or device is touched (dummy pointers)."""
    L = _native.lib()
    dummy = [ctypes.c_void_p(256)] * 8
    need = L.cbct_fp_backward_workspace_bytes_ex(64, 64, 64, 2000, 768, 640, None)
    assert need > _native.NOALLOC_SCRATCH_BYTES
    st = L.cbct_fp_backward_noalloc(ctypes.c_void_p(256), ctypes.c_void_p(256), 64, 64, 64, *dummy, 2000, 768, 640,
                                    1.0, 0.5, 1.0, 1, 0.0, 0.0, 0.0, ctypes.c_void_p(256), ctypes.c_void_p(256),
                                    None, None)
    assert st == 2 and b"fixed scratch" in L.cbct_last_error()

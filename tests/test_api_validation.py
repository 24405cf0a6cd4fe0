"""The public API's argument checks raise the reference's exception types and messages in the
reference's order (cbct_backprojector/operator.py:73-127, source_only_forward_projector.py:30-71,
filtering.py:17-51,139-179), with this library's own size limits; per-call settings are
per thread.  CPU only: the first check (float32 tensors on a GPU) is tested as it is, and for
the checks after it that check is replaced by a stub so that CPU tensors reach them; no call
reaches the library's compute entry points."""

import ctypes
import threading

import pytest
import torch

import cases  # noqa: F401  (sets sys.path)
from cbct_backprojector import _arguments, _native, backproject, detector_centers, fdk_filter, project
from cbct_backprojector import reconstruct_fdk


@pytest.fixture
def no_device_check(monkeypatch):
    monkeypatch.setattr(_arguments, "gpu_float32", lambda named, anchor: None)


def _geom(A):
    return [torch.zeros(A) for _ in range(8)]


def test_project_checks_in_reference_order(no_device_check):
    vol = torch.rand(3, 4, 5)
    g = _geom(2)
    with pytest.raises(ValueError, match=r"^volume must have a nonempty \[D,H,W\] shape$"):
        project(torch.rand(3, 4), *g, 1.0, 4, 4)
    with pytest.raises(ValueError, match=r"^volume must have a nonempty \[D,H,W\] shape$"):
        project(torch.rand(3, 0, 4), *g, 1.0, 4, 4)
    with pytest.raises(ValueError, match=r"^angles must have a nonempty shape \[A\]$"):
        project(vol, torch.zeros(0), *g[1:], 1.0, 4, 4)
    bad = list(g)
    bad[3] = torch.zeros(3)
    with pytest.raises(ValueError, match=r"^det_u0_arr must have shape \[2\]$"):
        project(vol, *bad, 1.0, 4, 4)
    with pytest.raises(TypeError, match=r"^detector_u must be a positive integer$"):
        project(vol, *g, 1.0, 4.0, 4)
    with pytest.raises(TypeError, match=r"^detector_v must be a positive integer$"):
        project(vol, *g, 1.0, 4, True)
    with pytest.raises(ValueError, match=r"^detector_v must be a positive integer$"):
        project(vol, *g, 1.0, 4, 0)
    with pytest.raises(TypeError, match=r"^off_y must be a finite Python number$"):
        project(vol, *g, 1.0, 4, 4, 0.5, 1.0, 0.0, "1", 0.0)
    with pytest.raises(ValueError, match=r"^off_z must be finite$"):
        project(vol, *g, 1.0, 4, 4, 0.5, 1.0, 0.0, 0.0, float("inf"))
    # the offsets are checked before the positive scalars
    with pytest.raises(ValueError, match=r"^off_x must be finite$"):
        project(vol, *g, -1.0, 4, 4, 0.5, 1.0, float("nan"))
    with pytest.raises(TypeError, match=r"^pixel_size must be a positive Python number$"):
        project(vol, *g, None, 4, 4)
    with pytest.raises(ValueError, match=r"^step_size must be finite and positive$"):
        project(vol, *g, 1.0, 4, 4, step_size=0.0)
    with pytest.raises(ValueError, match=r"^voxel_size must be finite and positive$"):
        project(vol, *g, 1.0, 4, 4, voxel_size=float("nan"))


def test_project_device_check():
    vol = torch.rand(3, 4, 5)
    with pytest.raises(TypeError, match=r"^volume must be a torch.Tensor$"):
        project(vol.numpy(), *_geom(2), 1.0, 4, 4)
    with pytest.raises(ValueError, match=r"^volume must be a CUDA float32 tensor$"):
        project(vol, *_geom(2), 1.0, 4, 4)


def test_backproject_checks_in_reference_order(no_device_check):
    p = torch.rand(2, 4, 4)
    g = _geom(2)
    with pytest.raises(ValueError, match=r"^projections must have shape \[A,U,V\]$"):
        backproject(torch.rand(2, 4), *g, 1.0, (2, 2, 2), 1.0)
    with pytest.raises(ValueError, match=r"^projections requires A >= 1 and U,V >= 2$"):
        backproject(torch.rand(2, 1, 4), *g, 1.0, (2, 2, 2), 1.0)
    bad = list(g)
    bad[7] = torch.zeros(2, 1)
    with pytest.raises(ValueError, match=r"^src_z must have shape \[2\]$"):
        backproject(p, *bad, 1.0, (2, 2, 2), 1.0)
    with pytest.raises(ValueError, match=r"^pixel_size must be finite and positive$"):
        backproject(p, *g, -1.0, (2, 2, 2), 1.0)
    with pytest.raises(TypeError, match=r"^voxel_size must be a positive Python number$"):
        backproject(p, *g, 1.0, (2, 2, 2), True)
    for shape in ((2, 2), (2, 2, 0), (2, 2.0, 2), (2, True, 2)):
        with pytest.raises(ValueError, match=r"^volume_shape must contain three positive integers \(D,H,W\)$"):
            backproject(p, *g, 1.0, shape, 1.0)
    with pytest.raises(ValueError, match=r"^Input or volume exceeds the kernel's 32-bit indexing limit$"):
        backproject(p, *g, 1.0, (2048, 1024, 1024), 1.0)
    # this library's launch limit (not the reference's CUDA grid: D, H up to 8 * 65535 pass)
    with pytest.raises(ValueError, match=r"^D exceeds the backprojector's launch limit of 4194240 voxels$"):
        backproject(p, *g, 1.0, (4194241, 1, 1), 1.0)
    for off in ((0.0, 0.0), (0.0, float("nan"), 0.0), (0.0, False, 0.0)):
        with pytest.raises(ValueError, match=r"^volume_offset must contain three finite numbers \(x,y,z\)$"):
            backproject(p, *g, 1.0, (2, 2, 2), 1.0, off)


def test_backproject_accepts_sizes_the_reference_grid_refused(no_device_check, monkeypatch):
    """D or H above 8 * 65535 is refused by the reference's CUDA grid check; this library's
    kernels have no such limit below its own launch limit, so validation passes and the call
    reaches the autograd node."""
    reached = []
    from cbct_backprojector import operator as op

    class Stub:
        @staticmethod
        def apply(*args):
            reached.append(args[10])

    monkeypatch.setattr(op, "_Backproject", Stub)
    backproject(torch.rand(2, 4, 4), *_geom(2), 1.0, (600000, 1, 1), 1.0)
    backproject(torch.rand(2, 4, 4), *_geom(2), 1.0, (1, 600000, 1), 1.0)
    assert reached == [(600000, 1, 1), (1, 600000, 1)]


def test_detector_centers_interface():
    u0, v0 = detector_centers(0.6, -0.3, 48, 40, 0.6)
    assert u0 == pytest.approx(23.5 - 1.0) and v0 == pytest.approx(19.5 + 0.5)
    with pytest.raises(ValueError, match=r"^detector_v must be an integer >= 2$"):
        detector_centers(0.0, 0.0, 4, 1, 1.0)
    with pytest.raises(ValueError, match=r"^detector_u must be an integer >= 2$"):
        detector_centers(0.0, 0.0, True, 4, 1.0)
    with pytest.raises(TypeError, match=r"^pixel_size must be a positive Python number$"):
        detector_centers(0.0, 0.0, 4, 4, "1")


def test_fdk_checks_in_reference_order():
    p = torch.rand(2, 4, 4)
    with pytest.raises(TypeError, match=r"^projections must be a torch.Tensor$"):
        fdk_filter(p.numpy(), 1.0, 100.0, 150.0)
    with pytest.raises(ValueError, match=r"^projections must have a nonempty \[A, U, V\] shape$"):
        fdk_filter(torch.rand(2, 0, 4), 1.0, 100.0, 150.0)
    with pytest.raises(TypeError, match=r"^projections must have dtype torch.float32$"):
        fdk_filter(p.double(), 1.0, 100.0, 150.0)
    with pytest.raises(ValueError, match=r"^pixel_size must be finite and positive$"):
        fdk_filter(p, 0.0, 100.0, 150.0)
    with pytest.raises(ValueError, match=r"^alpha must be finite and positive$"):
        fdk_filter(p, 1.0, 100.0, 150.0, alpha=0.0)
    with pytest.raises(ValueError, match=r"^alpha must satisfy 0 < alpha <= 1$"):
        fdk_filter(p, 1.0, 100.0, 150.0, alpha=1.5)
    with pytest.raises(ValueError, match=r"^sod must be a scalar or have one value per view$"):
        fdk_filter(p, 1.0, [1.0, 2.0, 3.0], 150.0)
    with pytest.raises(ValueError, match=r"^sid must contain finite positive distances$"):
        fdk_filter(p, 1.0, 100.0, [150.0, -1.0])
    with pytest.raises(ValueError, match="runs on the GPU only"):
        fdk_filter(p, 1.0, 100.0, 150.0)
    g = _geom(2)
    with pytest.raises(ValueError, match=r"^reconstruct_fdk requires CUDA projections$"):
        reconstruct_fdk(p, *g, 1.0, (2, 2, 2), 1.0)


def test_settings_are_per_thread():
    """Two threads select different kernels and deterministic modes; each reads back its own and
    builds its own per-call options, and neither changes the library's process-wide default."""
    L = _native.lib()
    default = ((1,) + (12, 5120, 0), (1,) + (12, 3840, 1))
    results = {}
    barrier = threading.Barrier(2)

    def worker(name, tuning, det):
        _native.set_fp_tuning(*tuning, op=_native.TUNING_SCATTER)
        _native.set_deterministic(det)
        barrier.wait()                         # both threads have set their settings
        snapshot = _native.CallSettings()
        barrier.wait()
        results[name] = (_native.get_fp_tuning(_native.TUNING_SCATTER), _native.get_deterministic(),
                         _native.options(_native.TUNING_SCATTER).as_tuple(),
                         snapshot.options(_native.TUNING_SCATTER).as_tuple(),
                         _native.get_fp_tuning(_native.TUNING_GATHER))

    t1 = threading.Thread(target=worker, args=("a", (1, 4, 2048, 0), True))
    t2 = threading.Thread(target=worker, args=("b", (0, 16, 8192, 1), False))
    t1.start(); t2.start(); t1.join(); t2.join()
    assert results["a"] == ((1, 4, 2048, 0), True, (1, 4, 2048, 0, 1), (1, 4, 2048, 0, 1), default[0])
    assert results["b"] == ((0, 16, 8192, 1), False, (0, 16, 8192, 1, 0), (0, 16, 8192, 1, 0), default[0])
    # this thread and the library's defaults are unchanged
    assert _native.get_fp_tuning(_native.TUNING_SCATTER) == default[1]
    out = _native.Options()
    assert L.cbct_fp_default_options(1, ctypes.byref(out)) == 0
    assert out.as_tuple() == default[1] + (0,)


def test_per_call_options_are_validated_by_the_library():
    """cbct_fp_forward_ex refuses out-of-range options before any launch (dummy non-null
    pointers are never dereferenced); NULL options mean the process-wide defaults."""
    L = _native.lib()
    dummy = [ctypes.c_void_p(256)] * 8
    bad = _native.Options(1, 12, 100, 0, 0)
    st = L.cbct_fp_forward_ex(ctypes.c_void_p(256), 8, 8, 8, *dummy, 2, 1.0, 16, 16, 0.5, 1.0, 0.0, 0.0, 0.0,
                              ctypes.c_void_p(256), ctypes.byref(bad), ctypes.c_void_p(256), 1 << 20, None)
    assert st == 1 and b"lds_floats" in L.cbct_last_error()
    bad = _native.Options(1, 12, 4096, 0, 2)
    st = L.cbct_fp_backward_ex(ctypes.c_void_p(256), ctypes.c_void_p(256), 8, 8, 8, *dummy, 2, 16, 16, 1.0, 0.5,
                               1.0, 1, 0.0, 0.0, 0.0, ctypes.c_void_p(256), None, ctypes.byref(bad),
                               ctypes.c_void_p(256), 1 << 20, None)
    assert st == 1 and b"deterministic" in L.cbct_last_error()
    # neither gradient requested
    ok = _native.Options(1, 12, 4096, 0, 0)
    st = L.cbct_fp_backward_ex(ctypes.c_void_p(256), ctypes.c_void_p(256), 8, 8, 8, *dummy, 2, 16, 16, 1.0, 0.5,
                               1.0, 0, 0.0, 0.0, 0.0, None, None, ctypes.byref(ok), ctypes.c_void_p(256), 1 << 20,
                               None)
    assert st == 1 and b"nothing to compute" in L.cbct_last_error()
    with pytest.raises(RuntimeError, match="lds_floats"):
        _native.options(tuning=(1, 12, 100, 0))

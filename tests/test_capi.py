"""The C-ABI library loads on the CPU host and exports every declared symbol;
the Python API validates its inputs exactly as the reference's wrappers do.
No compute calls (no GPU here)."""

import ctypes
import os
import re

import pytest
import torch

import cases  # noqa: F401  (sets sys.path)
from cbct_backprojector import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", name) for name in sorted(os.listdir(os.path.join(ROOT, "include")))
           if name.endswith(".h")]


def declared_symbols():
    text = "".join(open(h).read() for h in HEADERS)
    return sorted(set(re.findall(r"\b(cbct_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.LIBRARY_PATH)
    names = declared_symbols()
    assert len(names) >= 12
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_native.EXPORTED_SYMBOLS)


def test_version_and_workspace_queries():
    assert "gfx950" in _native.version()
    L = _native.lib()
    assert L.cbct_fp_workspace_bytes(720, 768, 640) > 720 * 64
    assert L.cbct_fp_workspace_bytes(0, 4, 4) == 0
    base = L.cbct_fp_backward_workspace_bytes(64, 64, 64, 10, 32, 32)
    assert base == L.cbct_fp_workspace_bytes(10, 32, 32)
    with _native.deterministic():
        opts = _native.options(_native.TUNING_SCATTER)
        det = L.cbct_fp_backward_workspace_bytes_ex(64, 64, 64, 10, 32, 32, ctypes.byref(opts))
    assert det == base + 64 ** 3 * 8 + 256
    assert not _native.get_deterministic()
    # the thread's setting never reaches the process-wide default
    assert L.cbct_fp_backward_workspace_bytes(64, 64, 64, 10, 32, 32) == base
    assert L.cbct_fp_get_deterministic() == 0
    assert L.cbct_bp_workspace_bytes(360, 256, 256, 256) == 360 * 64
    assert L.cbct_fdk_workspace_bytes(360, 512, 384) >= 360 * 512 * 384 * 4


def test_invalid_arguments_return_status_without_touching_the_device():
    """Argument checks run on the host before any launch (null pointers are never dereferenced)."""
    L = _native.lib()
    nul = [None] * 8
    st = L.cbct_fp_forward(None, 4, 4, 4, *nul, 3, 1.0, 4, 4, 0.5, 1.0, 0.0, 0.0, 0.0, None, None, 0, None)
    assert st != 0 and b"null" in L.cbct_last_error()
    st = L.cbct_fp_forward(None, 4, 4, 4, *nul, 3, -1.0, 4, 4, 0.5, 1.0, 0.0, 0.0, 0.0, None, None, 0, None)
    assert st == 1 and b"pixel_size" in L.cbct_last_error()
    st = L.cbct_bp_forward(None, 3, 1, 4, *nul, 1.0, 4, 4, 4, 1.0, 0.0, 0.0, 0.0, None, None, 0, None)
    assert st == 1 and b"U,V >= 2" in L.cbct_last_error()
    st = L.cbct_fdk_filter(None, 3, 4, 4, 1.0, None, None, 1.5, None, None, 0, None)
    assert st == 1 and b"alpha" in L.cbct_last_error()
    st = L.cbct_fdk_reconstruct(None, 3, 4, 4, *nul, 1.0, 4, 4, 4, 1.0, 0.0, 0.0, 0.0, 0.8, 0, None, None, 0, None)
    assert st == 1


def test_tuning_slots_set_get_and_validate():
    """The two tuning slots are independent host state; cbct_fp_set_tuning sets both and invalid
    settings are rejected with a status (no device involved)."""
    from cbct_backprojector import autotune
    old = (_native.get_fp_tuning(_native.TUNING_GATHER), _native.get_fp_tuning(_native.TUNING_SCATTER))
    try:
        assert old == ((1,) + autotune.DEFAULT_GATHER, (1,) + autotune.DEFAULT_SCATTER)
        _native.set_fp_tuning(1, 4, 2048, 1, op=_native.TUNING_GATHER)
        assert _native.get_fp_tuning(_native.TUNING_GATHER) == (1, 4, 2048, 1)
        assert _native.get_fp_tuning(_native.TUNING_SCATTER) == old[1]
        with _native.fp_tuning(0, 8, 4096, 0):
            assert _native.get_fp_tuning(0) == _native.get_fp_tuning(1) == (0, 8, 4096, 0)
        assert _native.get_fp_tuning(_native.TUNING_GATHER) == (1, 4, 2048, 1)
        L = _native.lib()
        assert L.cbct_fp_set_tuning_op(2, 1, 8, 4096, 1) == 1 and b"op must be" in L.cbct_last_error()
        assert L.cbct_fp_set_tuning_op(0, 1, 8, 100, 1) == 1 and b"lds_floats" in L.cbct_last_error()
        with pytest.raises(RuntimeError):
            _native.set_fp_tuning(1, 0, 4096, 1, op=1)
    finally:
        _native.set_fp_tuning(*old[0], op=_native.TUNING_GATHER)
        _native.set_fp_tuning(*old[1], op=_native.TUNING_SCATTER)


def _cpu_fp_args():
    vol = torch.rand(3, 4, 5)
    g = [torch.zeros(2) for _ in range(8)]
    return vol, g


def test_project_validation_errors_match_reference():
    from cbct_backprojector import project
    vol, g = _cpu_fp_args()
    with pytest.raises(ValueError, match="volume must be a CUDA float32 tensor"):
        project(vol, *g, 1.0, 4, 4)
    with pytest.raises(TypeError, match="volume must be a torch.Tensor"):
        project(vol.numpy(), *g, 1.0, 4, 4)


def test_backproject_validation_errors_match_reference():
    from cbct_backprojector import backproject
    p = torch.rand(2, 4, 4)
    g = [torch.zeros(2) for _ in range(8)]
    with pytest.raises(ValueError, match="projections must be a CUDA float32 tensor"):
        backproject(p, *g, 1.0, (2, 2, 2), 1.0)
    with pytest.raises(TypeError, match="projections must be a torch.Tensor"):
        backproject(p.numpy(), *g, 1.0, (2, 2, 2), 1.0)


def test_detector_centers_and_positive_scalar():
    from cbct_backprojector import detector_centers
    from cbct_backprojector.operator import _positive_scalar
    du = torch.tensor([0.3, -0.2], requires_grad=True)
    u0, v0 = detector_centers(du, 0.5, 48, 40, 0.8)
    assert torch.allclose(u0, torch.tensor([23.5 - 0.375, 23.5 + 0.25]))
    assert v0 == pytest.approx(19.5 - 0.625)
    u0.sum().backward()
    assert torch.allclose(du.grad, torch.full((2,), -1 / 0.8))
    with pytest.raises(ValueError, match="detector_u must be an integer >= 2"):
        detector_centers(0.0, 0.0, 1, 4, 1.0)
    with pytest.raises(TypeError):
        _positive_scalar(True, "x")
    with pytest.raises(ValueError):
        _positive_scalar(float("nan"), "x")
    with pytest.raises(ValueError):
        _positive_scalar(0.0, "x")


def test_fdk_validation_errors_match_reference():
    from cbct_backprojector import fdk_filter, reconstruct_fdk
    p = torch.rand(2, 4, 4)
    with pytest.raises(ValueError, match="alpha must satisfy"):
        fdk_filter(p, 1.0, 100.0, 150.0, alpha=1.5)
    with pytest.raises(ValueError, match="sod must be a scalar or have one value per view"):
        fdk_filter(p, 1.0, [1.0, 2.0, 3.0], 150.0)
    with pytest.raises(ValueError, match="runs on the GPU"):
        fdk_filter(p, 1.0, 100.0, 150.0)
    g = [torch.zeros(2) for _ in range(8)]
    with pytest.raises(ValueError, match="requires CUDA projections"):
        reconstruct_fdk(p, *g, 1.0, (2, 2, 2), 1.0)


def test_compat_module_paths():
    import source_only_forward_projector as shim
    import cbct_backprojector
    assert shim.project is cbct_backprojector.project
    import src_cbct_backprojector_cuda
    import src_cbct_forwardprojector_cuda
    assert callable(src_cbct_forwardprojector_cuda.forward) and callable(src_cbct_backprojector_cuda.backward)
    assert cbct_backprojector.__version__ == "0.1.0"

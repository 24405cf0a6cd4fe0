"""ctypes binding of libcbct_hip.so, the gfx950 operator library.

The library has a plain C ABI (``include/cbct_hip.h``): device pointers, int64
sizes, float scalars, a caller-owned workspace and a stream.  This module
allocates outputs and workspaces with torch on the input's device, passes
torch's current stream, and turns a non-zero status into ``RuntimeError`` with
the library's message.  There is no other implementation behind it: when the
library is missing every operator raises.
"""

import ctypes
import os
import threading

import torch

_LIB = None
_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "libcbct_hip.so")

_BUILD_HELP = (
    "The HIP operator library is not built ({path} is missing). From the repository "
    "root run: make -C diff-geocali_amd/csrc  (or python -c \"import __graft_entry__ as g; g.build()\")"
)

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_F = ctypes.c_float
_SZ = ctypes.c_size_t
_INT = ctypes.c_int

_SIGNATURES = {
    "cbct_version": ([], ctypes.c_char_p),
    "cbct_last_error": ([], ctypes.c_char_p),
    "cbct_fp_workspace_bytes": ([_I64, _I64, _I64], _SZ),
    "cbct_fp_forward": ([_P, _I64, _I64, _I64] + [_P] * 8 + [_I64, _F, _I64, _I64, _F, _F, _F, _F, _F,
                        _P, _P, _SZ, _P], _INT),
    "cbct_fp_backward": ([_P, _P, _I64, _I64, _I64] + [_P] * 8 + [_I64, _I64, _I64, _F, _F, _F, _INT,
                         _F, _F, _F, _P, _P, _P, _SZ, _P], _INT),
    "cbct_fp_count_samples": ([_I64, _I64, _I64] + [_P] * 8 + [_I64, _F, _I64, _I64, _F, _F, _F, _F, _F,
                              _P, _P, _SZ, _P], _INT),
    "cbct_fp_set_tuning": ([_INT, _INT, _INT, _INT], _INT),
    "cbct_fp_backward_workspace_bytes": ([_I64, _I64, _I64, _I64, _I64, _I64], _SZ),
    "cbct_fp_set_deterministic": ([_INT], _INT),
    "cbct_fp_get_deterministic": ([], _INT),
    "cbct_fp_get_tuning": ([_P], _INT),
    "cbct_fp_set_tuning_op": ([_INT, _INT, _INT, _INT, _INT], _INT),
    "cbct_fp_get_tuning_op": ([_INT, _P], _INT),
    "cbct_fp_default_options": ([_INT, _P], _INT),
    "cbct_fp_forward_ex": ([_P, _I64, _I64, _I64] + [_P] * 8 + [_I64, _F, _I64, _I64, _F, _F, _F, _F, _F,
                           _P, _P, _P, _SZ, _P], _INT),
    "cbct_fp_backward_workspace_bytes_ex": ([_I64] * 6 + [_P], _SZ),
    "cbct_fp_backward_ex": ([_P, _P, _I64, _I64, _I64] + [_P] * 8 + [_I64, _I64, _I64, _F, _F, _F, _INT,
                            _F, _F, _F, _P, _P, _P, _P, _SZ, _P], _INT),
    "cbct_fp_backward_noalloc": ([_P, _P, _I64, _I64, _I64] + [_P] * 8 + [_I64, _I64, _I64, _F, _F, _F, _INT,
                                 _F, _F, _F, _P, _P, _P, _P], _INT),
    "cbct_bp_backward_workspace_bytes_ex": ([_I64] * 6 + [_P], _SZ),
    "cbct_bp_backward_ex": ([_P, _P, _I64, _I64, _I64] + [_P] * 8 + [_F, _I64, _I64, _I64, _F, _F, _F, _F,
                            _P, _P, _P, _P, _SZ, _P], _INT),
    "cbct_bp_workspace_bytes": ([_I64, _I64, _I64, _I64], _SZ),
    "cbct_bp_backward_workspace_bytes": ([_I64] * 6, _SZ),
    "cbct_bp_forward": ([_P, _I64, _I64, _I64] + [_P] * 8 + [_F, _I64, _I64, _I64, _F, _F, _F, _F,
                        _P, _P, _SZ, _P], _INT),
    "cbct_bp_backward": ([_P, _P, _I64, _I64, _I64] + [_P] * 8 + [_F, _I64, _I64, _I64, _F, _F, _F, _F,
                         _P, _P, _P, _SZ, _P], _INT),
    "cbct_fdk_workspace_bytes": ([_I64, _I64, _I64], _SZ),
    "cbct_fdk_filter": ([_P, _I64, _I64, _I64, _F, _P, _P, _F, _P, _P, _SZ, _P], _INT),
    "cbct_fdk_reconstruct": ([_P, _I64, _I64, _I64] + [_P] * 8 + [_F, _I64, _I64, _I64, _F, _F, _F, _F,
                             _F, _I64, _P, _P, _SZ, _P], _INT),
    "cbct_fdk_filter_ex": ([_P, _I64, _I64, _I64, _F, _P, _P, _F, _P, _P, _P, _P, _P, _SZ, _P], _INT),
    "cbct_fdk_reconstruct_ex": ([_P, _I64, _I64, _I64] + [_P] * 8 + [_F, _I64, _I64, _I64, _F, _F, _F, _F,
                                _F, _I64, _INT, _INT, _P, _P, _SZ, _P], _INT),
    "cbct_calib_workspace_bytes": ([_P, _INT], _SZ),
    "cbct_calib_run": ([_P, _INT, _INT, _INT, _P, _F, _F, _F, _F, _INT, _P, _SZ, _P], _INT),
    "cbct_calib_lm_workspace_bytes": ([_P, _INT], _SZ),
    "cbct_calib_run_lm": ([_P, _INT, _INT, _F, _INT, _P, _SZ, _P], _INT),
    "cbct_calib_release_graphs": ([], _INT),
    "cbct_preprocess": ([_P, _INT, _P, _P, _I64, _I64, _I64, _INT, _F, _INT, _P, _P], _INT),
    # learned warm start (include/cbct_nn.h)
    "cbct_nn_gemm_workspace_bytes": ([_I64, _I64, _INT], _SZ),
    "cbct_nn_gemm": ([_INT, _INT, _I64, _I64, _I64, _P, _I64, _P, _I64, _P, _INT, _P, _I64, _P, _I64, _INT,
                      _P, _SZ, _P], _INT),
    "cbct_nn_im2col": ([_P, _I64, _I64, _I64, _I64, _INT, _I64, _P, _P], _INT),
    "cbct_nn_col2im": ([_P, _I64, _I64, _I64, _I64, _INT, _I64, _P, _P, _P], _INT),
    "cbct_nn_relu_backward": ([_P, _P, _I64, _P, _P, _P], _INT),
    "cbct_nn_spatial_mean": ([_P, _I64, _I64, _I64, _P, _P], _INT),
    "cbct_nn_spatial_mean_backward": ([_P, _I64, _I64, _I64, _P, _P, _P], _INT),
    "cbct_nn_view_concat": ([_P, _I64, _I64, _I64, _P, _P], _INT),
    "cbct_nn_view_concat_backward": ([_P, _I64, _I64, _I64, _P, _P], _INT),
    "cbct_nn_colsum": ([_P, _I64, _I64, _P, _P], _INT),
    "cbct_nn_mse": ([_P, _P, _I64, _P, _P, _P], _INT),
    "cbct_nn_adam": ([_P, _P, _P, _P, _P, _I64, _F, _F, _F, _F, _F, _F, _F, _P], _INT),
    "cbct_nn_to_bf16": ([_P, _I64, _P, _P], _INT),
    "cbct_nn_prepare_input": ([_P, _P, _I64, _I64, _I64, _INT, _F, _F, _P, _P], _INT),
    "cbct_nn_jacobian_mse": ([_P, _P, _P, _I64, _I64, _I64, _I64, _P, _P, _P], _INT),
    "cbct_nn_sqdiff_workspace_bytes": ([], _SZ),
    "cbct_nn_sqdiff_mean": ([_P, _P, _I64, _P, _P, _SZ, _P], _INT),
}

EXPORTED_SYMBOLS = tuple(_SIGNATURES)


def lib():
    """Load (once) and return the operator library; raise RuntimeError when it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBRARY_PATH):
            raise RuntimeError(_BUILD_HELP.format(path=LIBRARY_PATH))
        handle = ctypes.CDLL(LIBRARY_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = handle
    return _LIB


def version():
    return lib().cbct_version().decode()


def _check(status):
    if status != 0:
        raise RuntimeError(lib().cbct_last_error().decode())


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _workspace(nbytes, device):
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def _require_gpu(t, what):
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
        raise RuntimeError(f"{what} must be a GPU tensor (ROCm devices appear as 'cuda')")


# Tuning slots: 0 = kernels that only gather (forward, geometry-only adjoint, calibration march),
# 1 = the adjoint with the volume gradient.
TUNING_GATHER, TUNING_SCATTER = 0, 1


class Options(ctypes.Structure):
    """cbct_fp_options: the kernel selection and deterministic mode of one call."""
    _fields_ = [("use_slab", _INT), ("slab_thickness", _INT), ("lds_floats", _INT), ("layout", _INT),
                ("deterministic", _INT)]

    def as_tuple(self):
        return (self.use_slab, self.slab_thickness, self.lds_floats, self.layout, self.deterministic)


def _validate_tuning(use_slab, slab_thickness, lds_floats, layout):
    """The library's ranges (cbct_fp_options), checked before any setting is kept."""
    if use_slab not in (0, 1):
        raise RuntimeError("use_slab must be 0 or 1")
    if not 1 <= slab_thickness <= 64:
        raise RuntimeError("slab_thickness must be in [1, 64]")
    if not 256 <= lds_floats <= 8192:
        raise RuntimeError("lds_floats must be in [256, 8192]")
    if layout not in (0, 1):
        raise RuntimeError("layout must be 0 or 1")


# Per-thread settings.  The library's own process-wide defaults are read once per thread;
# set_fp_tuning / set_deterministic (and their context managers) then change only the calling
# thread's settings, and every call passes them to the library explicitly (cbct_fp_options), so
# threads with different settings never see each other's.  The autograd Functions capture the
# settings when their forward runs, so a backward run by autograd's own threads uses them too.
_THREAD = threading.local()


def _state():
    st = getattr(_THREAD, "state", None)
    if st is None:
        L = lib()
        slots = []
        for op in (TUNING_GATHER, TUNING_SCATTER):
            o = Options()
            _check(L.cbct_fp_default_options(op, ctypes.byref(o)))
            slots.append(o.as_tuple()[:4])
        st = {"tuning": slots, "deterministic": bool(o.deterministic)}
        _THREAD.state = st
    return st


def set_fp_tuning(use_slab, slab_thickness, lds_floats, layout, op=None):
    """Select the projector kernels for this thread's later calls (same results for every setting),
    for one tuning slot (``op`` 0 or 1) or both (``op=None``)."""
    args = (int(use_slab), int(slab_thickness), int(lds_floats), int(layout))
    _validate_tuning(*args)
    st = _state()
    for k in ((TUNING_GATHER, TUNING_SCATTER) if op is None else (int(op),)):
        if k not in (TUNING_GATHER, TUNING_SCATTER):
            raise RuntimeError("op must be 0 (gathering kernels) or 1 (volume-gradient adjoint)")
        st["tuning"][k] = args


def get_fp_tuning(op=TUNING_GATHER):
    return tuple(_state()["tuning"][int(op)])


def set_deterministic(on):
    """Bitwise-reproducible projector volume gradients and backprojection projection gradients
    (int64 fixed-point accumulation) for this thread's later calls."""
    _state()["deterministic"] = bool(on)


def get_deterministic():
    return bool(_state()["deterministic"])


def options(op=TUNING_GATHER, tuning=None, deterministic=None):
    """cbct_fp_options of one call in slot ``op``: this thread's settings, or the given
    ``tuning`` (use_slab, slab_thickness, lds_floats, layout) / ``deterministic``."""
    st = _state()
    t = tuple(int(x) for x in (tuning if tuning is not None else st["tuning"][int(op)]))
    _validate_tuning(*t)
    det = st["deterministic"] if deterministic is None else bool(deterministic)
    return Options(*t, 1 if det else 0)


class CallSettings:
    """A snapshot of this thread's settings (both slots and the deterministic mode), taken by an
    autograd Function's forward and used again by its backward."""

    def __init__(self):
        st = _state()
        self.tuning = [tuple(x) for x in st["tuning"]]
        self.deterministic = st["deterministic"]

    def options(self, op):
        return options(op, tuning=self.tuning[op], deterministic=self.deterministic)


def set_default_fp_tuning(use_slab, slab_thickness, lds_floats, layout, op=None):
    """Change the library's process-wide default (read by threads that have not read it yet and by
    C callers that pass no options)."""
    args = (int(use_slab), int(slab_thickness), int(lds_floats), int(layout))
    if op is None:
        _check(lib().cbct_fp_set_tuning(*args))
    else:
        _check(lib().cbct_fp_set_tuning_op(int(op), *args))


class deterministic:
    """Context manager for this thread's deterministic mode."""

    def __init__(self, on=True):
        self.new = bool(on)

    def __enter__(self):
        self.old = get_deterministic()
        set_deterministic(self.new)
        return self

    def __exit__(self, *exc):
        set_deterministic(self.old)
        return False


class fp_tuning:
    """Context manager that applies a projector tuning to this thread (one slot, or both with
    ``op=None``) and restores both slots afterwards."""

    def __init__(self, use_slab, slab_thickness=8, lds_floats=4096, layout=1, op=None):
        self.new = (use_slab, slab_thickness, lds_floats, layout)
        self.op = op

    def __enter__(self):
        self.old = (get_fp_tuning(TUNING_GATHER), get_fp_tuning(TUNING_SCATTER))
        set_fp_tuning(*self.new, op=self.op)
        return self

    def __exit__(self, *exc):
        set_fp_tuning(*self.old[0], op=TUNING_GATHER)
        set_fp_tuning(*self.old[1], op=TUNING_SCATTER)
        return False


def fp_forward(volume, geometry, pixel_size, U, V, step_size, voxel_size, off_x, off_y, off_z, opts=None):
    """proj[A,U,V]; volume [D,H,W] and the 8 geometry tensors must be contiguous float32.
    ``opts``: cbct_fp_options of the call (default: this thread's gather settings)."""
    _require_gpu(volume, "volume")
    L = lib()
    D, H, W = volume.shape
    A = geometry[0].numel()
    dev = volume.device
    o = opts if opts is not None else options(TUNING_GATHER)
    with torch.cuda.device(dev):
        out = torch.empty((A, U, V), dtype=torch.float32, device=dev)
        ws_bytes = L.cbct_fp_workspace_bytes(A, U, V)
        ws = _workspace(ws_bytes, dev)
        _check(L.cbct_fp_forward_ex(_ptr(volume), D, H, W, *[_ptr(g) for g in geometry], A,
                                    pixel_size, U, V, step_size, voxel_size, off_x, off_y, off_z,
                                    _ptr(out), ctypes.byref(o), _ptr(ws), ws_bytes, _stream(dev)))
    return out


def fp_backward(grad, volume, geometry, pixel_size, step_size, voxel_size, compute_grad_volume,
                off_x, off_y, off_z, compute_grad_geometry=True, opts=None):
    """(gvol [D,H,W] or empty, [8, A] geometry gradients or None) of sum(grad * fp_forward).
    ``opts``: cbct_fp_options of the call (default: this thread's settings of the slot the call
    runs in)."""
    _require_gpu(volume, "volume")
    L = lib()
    D, H, W = volume.shape
    A, U, V = grad.shape
    dev = volume.device
    o = opts if opts is not None else options(TUNING_SCATTER if compute_grad_volume else TUNING_GATHER)
    with torch.cuda.device(dev):
        gvol = (torch.empty((D, H, W), dtype=torch.float32, device=dev) if compute_grad_volume
                else torch.empty((0,), dtype=torch.float32, device=dev))
        ggeo = torch.empty((8, A), dtype=torch.float32, device=dev) if compute_grad_geometry else None
        ws_bytes = L.cbct_fp_backward_workspace_bytes_ex(D, H, W, A, U, V, ctypes.byref(o))
        ws = _workspace(ws_bytes, dev)
        _check(L.cbct_fp_backward_ex(_ptr(grad), _ptr(volume), D, H, W, *[_ptr(g) for g in geometry],
                                     A, U, V, pixel_size, step_size, voxel_size,
                                     1 if compute_grad_volume else 0, off_x, off_y, off_z,
                                     _ptr(gvol) if compute_grad_volume else None, _ptr(ggeo), ctypes.byref(o),
                                     _ptr(ws), ws_bytes, _stream(dev)))
    return gvol, ggeo


# cbct_fp_backward_noalloc's fixed device scratch (CBCT_NOALLOC_SCRATCH_BYTES in cbct_hip.h)
NOALLOC_SCRATCH_BYTES = 64 << 20


def fp_backward_noalloc(grad, volume, geometry, pixel_size, step_size, voxel_size, compute_grad_volume,
                        off_x, off_y, off_z, grad_volume=None, grad_geometry=None, opts=None):
    """fp_backward through the allocation-free entry point: the library allocates nothing (its
    workspace is the fixed device scratch); the caller passes the outputs (``grad_volume``
    [D,H,W] when the volume gradient is wanted, ``grad_geometry`` [8, A] or None)."""
    _require_gpu(volume, "volume")
    L = lib()
    D, H, W = volume.shape
    A, U, V = grad.shape
    dev = volume.device
    o = opts if opts is not None else options(TUNING_SCATTER if compute_grad_volume else TUNING_GATHER)
    with torch.cuda.device(dev):
        _check(L.cbct_fp_backward_noalloc(_ptr(grad), _ptr(volume), D, H, W, *[_ptr(g) for g in geometry],
                                          A, U, V, pixel_size, step_size, voxel_size,
                                          1 if compute_grad_volume else 0, off_x, off_y, off_z,
                                          _ptr(grad_volume), _ptr(grad_geometry), ctypes.byref(o), _stream(dev)))
    return grad_volume, grad_geometry


def fp_count_samples(volume_shape, geometry, pixel_size, U, V, step_size, voxel_size,
                     off_x=0.0, off_y=0.0, off_z=0.0):
    """Exact number of trilinear samples the forward schedule takes (int)."""
    dev = geometry[0].device
    _require_gpu(geometry[0], "geometry")
    L = lib()
    D, H, W = volume_shape
    A = geometry[0].numel()
    with torch.cuda.device(dev):
        count = torch.zeros(1, dtype=torch.int64, device=dev)
        ws_bytes = L.cbct_fp_workspace_bytes(A, U, V)
        ws = _workspace(ws_bytes, dev)
        _check(L.cbct_fp_count_samples(D, H, W, *[_ptr(g) for g in geometry], A, pixel_size, U, V,
                                       step_size, voxel_size, off_x, off_y, off_z, _ptr(count),
                                       _ptr(ws), ws_bytes, _stream(dev)))
    return int(count.item())


def bp_forward(proj, geometry, pixel_size, D, H, W, voxel_size, off_x, off_y, off_z):
    _require_gpu(proj, "projections")
    L = lib()
    A, U, V = proj.shape
    dev = proj.device
    with torch.cuda.device(dev):
        vol = torch.empty((D, H, W), dtype=torch.float32, device=dev)
        ws_bytes = L.cbct_bp_workspace_bytes(A, D, H, W)
        ws = _workspace(ws_bytes, dev)
        _check(L.cbct_bp_forward(_ptr(proj), A, U, V, *[_ptr(g) for g in geometry], pixel_size,
                                 D, H, W, voxel_size, off_x, off_y, off_z, _ptr(vol), _ptr(ws),
                                 ws_bytes, _stream(dev)))
    return vol


def bp_backward(grad_vol, proj, geometry, pixel_size, voxel_size, off_x, off_y, off_z, opts=None):
    """(grad_projections [A,U,V], [8, A] geometry gradients); ``opts``: cbct_fp_options of the
    call (only its deterministic flag matters; default: this thread's setting)."""
    _require_gpu(proj, "projections")
    L = lib()
    A, U, V = proj.shape
    D, H, W = grad_vol.shape
    dev = proj.device
    o = opts if opts is not None else options(TUNING_GATHER)
    with torch.cuda.device(dev):
        gproj = torch.empty((A, U, V), dtype=torch.float32, device=dev)
        ggeo = torch.empty((8, A), dtype=torch.float32, device=dev)
        ws_bytes = L.cbct_bp_backward_workspace_bytes_ex(A, U, V, D, H, W, ctypes.byref(o))
        ws = _workspace(ws_bytes, dev)
        _check(L.cbct_bp_backward_ex(_ptr(grad_vol), _ptr(proj), A, U, V, *[_ptr(g) for g in geometry],
                                     pixel_size, D, H, W, voxel_size, off_x, off_y, off_z,
                                     _ptr(gproj), _ptr(ggeo), ctypes.byref(o), _ptr(ws), ws_bytes, _stream(dev)))
    return gproj, ggeo


def fdk_filter(proj, pixel_size, sod, sid, alpha, det_u0=None, det_v0=None, angles=None):
    """sod, sid: contiguous float32 [A] tensors on the projections' device; det_u0/det_v0
    (both or neither): per-view preweight centres; angles: Parker short-scan weights."""
    _require_gpu(proj, "projections")
    L = lib()
    A, U, V = proj.shape
    dev = proj.device
    opt = lambda t: None if t is None else _ptr(t)
    with torch.cuda.device(dev):
        out = torch.empty((A, U, V), dtype=torch.float32, device=dev)
        ws_bytes = L.cbct_fdk_workspace_bytes(A, U, V)
        ws = _workspace(ws_bytes, dev)
        _check(L.cbct_fdk_filter_ex(_ptr(proj), A, U, V, pixel_size, _ptr(sod), _ptr(sid), alpha, opt(det_u0),
                                    opt(det_v0), opt(angles), _ptr(out), _ptr(ws), ws_bytes, _stream(dev)))
    return out


def fdk_reconstruct(proj, geometry, pixel_size, D, H, W, voxel_size, off_x, off_y, off_z, alpha,
                    angle_chunk_size, short_scan=False, offset_preweight=False):
    _require_gpu(proj, "projections")
    L = lib()
    A, U, V = proj.shape
    dev = proj.device
    with torch.cuda.device(dev):
        vol = torch.empty((D, H, W), dtype=torch.float32, device=dev)
        ws_bytes = L.cbct_fdk_workspace_bytes(A, U, V)
        ws = _workspace(ws_bytes, dev)
        _check(L.cbct_fdk_reconstruct_ex(_ptr(proj), A, U, V, *[_ptr(g) for g in geometry], pixel_size,
                                         D, H, W, voxel_size, off_x, off_y, off_z, alpha, angle_chunk_size,
                                         1 if short_scan else 0, 1 if offset_preweight else 0,
                                         _ptr(vol), _ptr(ws), ws_bytes, _stream(dev)))
    return vol

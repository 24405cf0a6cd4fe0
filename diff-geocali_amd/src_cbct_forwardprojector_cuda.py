"""Low-level forward-projector module with the reference's native signatures.

Drop-in for the pybind module ``src_cbct_forwardprojector_cuda``
(csrc/src_cbct_forwardprojector_cuda.cu:732-757): ``forward`` and ``backward``
with the same argument names, defaults and return values, implemented over the
C ABI of libcbct_hip.so.  As in the reference, calling these directly does not
register autograd; use ``cbct_backprojector.project`` for that.
"""

import torch

from cbct_backprojector import _native

_GEOMETRY = ("angles", "sid", "sod", "det_u0", "det_v0", "src_x", "src_y", "src_z")


def _check_geometry(reference, tensors, expected):
    """check_geometry_inputs, csrc/src_cbct_forwardprojector_cuda.cu:531-563."""
    if reference.device.type != "cuda":
        raise RuntimeError("reference tensor must be CUDA")
    if reference.dtype != torch.float32:
        raise RuntimeError("reference tensor must be float32")
    for name, t in zip(_GEOMETRY, tensors):
        if t.device.type != "cuda":
            raise RuntimeError(f"{name} must be CUDA")
        if t.dtype != torch.float32:
            raise RuntimeError(f"{name} must be float32")
        if t.device != reference.device:
            raise RuntimeError(f"{name} must be on the same CUDA device as the reference tensor")
        if t.numel() != expected:
            raise RuntimeError(f"{name} must contain A elements")


def forward(volume, angles, sid, sod, det_u0_arr, det_v0_arr, src_x_arr, src_y_arr, src_z_arr,
            pixel_size, U, V, step_size, voxel_size, off_x=0.0, off_y=0.0, off_z=0.0):
    """Source-only cone-beam forward projection -> [A,U,V] (forward_cuda, :566-631)."""
    if volume.dim() != 3:
        raise RuntimeError("volume must have shape [D,H,W]")
    if not (U > 0 and V > 0):
        raise RuntimeError("U and V must be positive")
    if not (pixel_size > 0 and step_size > 0 and voxel_size > 0):
        raise RuntimeError("pixel_size, step_size and voxel_size must be positive")
    geometry = (angles, sid, sod, det_u0_arr, det_v0_arr, src_x_arr, src_y_arr, src_z_arr)
    A = angles.numel()
    if A <= 0:
        raise RuntimeError("angles must contain at least one view")
    _check_geometry(volume, geometry, A)
    if min(volume.shape) <= 0:
        raise RuntimeError("every volume dimension must be positive")
    return _native.fp_forward(volume.contiguous(), [g.contiguous() for g in geometry],
                              float(pixel_size), int(U), int(V), float(step_size),
                              float(voxel_size), float(off_x), float(off_y), float(off_z))


def backward(grad_projection, volume, angles, sid, sod, det_u0_arr, det_v0_arr,
             src_x_arr, src_y_arr, src_z_arr, pixel_size, step_size, voxel_size,
             compute_grad_volume=True, off_x=0.0, off_y=0.0, off_z=0.0):
    """Nine gradients [volume, angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z]
    (backward_cuda, :634-729); the volume gradient is an empty tensor when
    compute_grad_volume is false."""
    if volume.dim() != 3:
        raise RuntimeError("volume must have shape [D,H,W]")
    if grad_projection.dim() != 3:
        raise RuntimeError("grad_projection must have shape [A,U,V]")
    if grad_projection.device.type != "cuda" or grad_projection.dtype != torch.float32:
        raise RuntimeError("grad_projection must be a float32 CUDA tensor")
    if grad_projection.device != volume.device:
        raise RuntimeError("grad_projection and volume must be on the same CUDA device")
    if not (pixel_size > 0 and step_size > 0 and voxel_size > 0):
        raise RuntimeError("pixel_size, step_size and voxel_size must be positive")
    A, U, V = grad_projection.shape
    if not (A > 0 and U > 0 and V > 0):
        raise RuntimeError("every grad_projection dimension must be positive")
    geometry = (angles, sid, sod, det_u0_arr, det_v0_arr, src_x_arr, src_y_arr, src_z_arr)
    _check_geometry(volume, geometry, A)
    gvol, ggeo = _native.fp_backward(grad_projection.contiguous(), volume.contiguous(),
                                     [g.contiguous() for g in geometry], float(pixel_size),
                                     float(step_size), float(voxel_size), bool(compute_grad_volume),
                                     float(off_x), float(off_y), float(off_z))
    return [gvol] + [ggeo[k].reshape(geometry[k].shape).clone() for k in range(8)]

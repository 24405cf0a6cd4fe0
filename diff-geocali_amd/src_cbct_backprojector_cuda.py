"""Low-level backprojector module with the reference's native signatures.

Drop-in for the pybind module ``src_cbct_backprojector_cuda``
(csrc/src_cbct_backprojector_cuda.cu:524-543): ``forward`` and ``backward``
with the same argument names, defaults and return values, implemented over the
C ABI of libcbct_hip.so.  Calling these directly does not register autograd;
use ``cbct_backprojector.backproject`` for that.
"""

import torch

from cbct_backprojector import _native

_INT_MAX = 2**31 - 1


def _check_inputs(proj, geometry):
    """check_geometry_inputs, csrc/src_cbct_backprojector_cuda.cu:298-331."""
    if proj.device.type != "cuda" or proj.dtype != torch.float32:
        raise RuntimeError("proj must be a CUDA float32 tensor")
    if not (proj.dim() == 3 and proj.shape[0] > 0 and proj.shape[1] >= 2 and proj.shape[2] >= 2):
        raise RuntimeError("proj must have shape [A,U,V] with A >= 1 and U,V >= 2")
    if proj.numel() > _INT_MAX:
        raise RuntimeError("proj exceeds the kernel's 32-bit indexing limit")
    names = ("angles", "sid", "sod", "det_u0", "det_v0", "src_x", "src_y", "src_z")
    for name, t in zip(names, geometry):
        if not (t.device.type == "cuda" and t.dtype == torch.float32 and t.device == proj.device
                and t.dim() == 1 and t.numel() == proj.shape[0]):
            raise RuntimeError(f"{name} must be CUDA float32 on proj.device with shape [A]")


def _check_volume_shape(D, H, W):
    """check_volume_shape, csrc/src_cbct_backprojector_cuda.cu:333-341."""
    if not (0 < D <= _INT_MAX and 0 < H <= _INT_MAX and 0 < W <= _INT_MAX):
        raise RuntimeError("D,H,W must be positive and fit in a 32-bit integer")
    if not (D <= _INT_MAX // H and D * H <= _INT_MAX // W):
        raise RuntimeError("volume exceeds the kernel's 32-bit indexing limit")
    if not (D <= 8 * 65535 and H <= 8 * 65535):
        raise RuntimeError("D,H exceed the kernel's 3D CUDA grid limits")


def _check_scalars(pixel_size, voxel_size, offsets):
    import math
    if not all(math.isfinite(x) for x in (pixel_size, voxel_size, *offsets)):
        raise RuntimeError("sizes and ROI offsets must be finite")
    if not (pixel_size > 0 and voxel_size > 0):
        raise RuntimeError("pixel_size and voxel_size must be positive")


def forward(proj, angles, sid, sod, det_u0_arr, det_v0_arr, src_x_arr, src_y_arr, src_z_arr,
            pixel_size, D, H, W, voxel_size, off_x=0.0, off_y=0.0, off_z=0.0):
    """Weighted backprojection -> [D,H,W] (forward_cuda, :344-420)."""
    geometry = (angles, sid, sod, det_u0_arr, det_v0_arr, src_x_arr, src_y_arr, src_z_arr)
    _check_inputs(proj, geometry)
    _check_volume_shape(D, H, W)
    _check_scalars(pixel_size, voxel_size, (off_x, off_y, off_z))
    return _native.bp_forward(proj.contiguous(), [g.contiguous() for g in geometry],
                              float(pixel_size), int(D), int(H), int(W), float(voxel_size),
                              float(off_x), float(off_y), float(off_z))


def backward(grad_vol, proj, angles, sid, sod, det_u0_arr, det_v0_arr,
             src_x_arr, src_y_arr, src_z_arr, pixel_size, voxel_size,
             off_x=0.0, off_y=0.0, off_z=0.0):
    """Nine gradients [proj, angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z]
    (backward_cuda, :423-522)."""
    geometry = (angles, sid, sod, det_u0_arr, det_v0_arr, src_x_arr, src_y_arr, src_z_arr)
    _check_inputs(proj, geometry)
    if not (grad_vol.device.type == "cuda" and grad_vol.dtype == torch.float32
            and grad_vol.device == proj.device and grad_vol.dim() == 3):
        raise RuntimeError("grad_vol must be CUDA float32 on proj.device with shape [D,H,W]")
    _check_volume_shape(*grad_vol.shape)
    _check_scalars(pixel_size, voxel_size, (off_x, off_y, off_z))
    gproj, ggeo = _native.bp_backward(grad_vol.contiguous(), proj.contiguous(),
                                      [g.contiguous() for g in geometry], float(pixel_size),
                                      float(voxel_size), float(off_x), float(off_y), float(off_z))
    return [gproj] + [ggeo[k].clone() for k in range(8)]

"""Weighted cone-beam backprojection (with first-order autograd) and detector centres.

Public names, signatures and errors are those of the reference package's
``cbct_backprojector/operator.py`` (``_positive_scalar``, ``detector_centers``,
``_Backproject``, ``backproject``); the arithmetic is ``cbct_bp_forward`` /
``cbct_bp_backward_ex`` of ``libcbct_hip.so``.
"""

from numbers import Integral

import torch
from torch.autograd.function import once_differentiable

from . import _arguments, _native

_GEOMETRY = ("angles", "sid", "sod", "det_u0", "det_v0", "src_x", "src_y", "src_z")


def _positive_scalar(value, name):
    """Return ``value`` as a float; refuse bools, non-real values and anything not finite and
    strictly positive (interface of cbct_backprojector/operator.py:11-17)."""
    return _arguments.positive_number(value, name)


def detector_centers(delta_u, delta_v, detector_u, detector_v, pixel_size):
    """Pixel indices (u0, v0) hit by the central ray, from physical detector shifts.

    The detector's middle is ((U - 1)/2, (V - 1)/2); shifting the detector by +delta (length
    units) along u or v moves the index of the central ray by -delta/pixel_size.  Plain numbers
    give plain numbers; tensors give tensors that keep their autograd history, so shifts can be
    fitted through this function (interface of cbct_backprojector/operator.py:32-45).
    """
    pixel = _positive_scalar(pixel_size, "pixel_size")
    for label, extent in (("detector_u", detector_u), ("detector_v", detector_v)):
        if isinstance(extent, bool) or not isinstance(extent, Integral) or extent < 2:
            raise ValueError(f"{label} must be an integer >= 2")
    middle_u = (detector_u - 1.0) / 2.0
    middle_v = (detector_v - 1.0) / 2.0
    return middle_u - delta_u / pixel, middle_v - delta_v / pixel


class _Backproject(torch.autograd.Function):
    """Autograd node of the backprojector: positional arguments as in the reference's
    ``_Backproject.apply`` (cbct_backprojector/operator.py:48-70); gradients reach the
    projections and the eight geometry tensors, never the scalar settings."""

    @staticmethod
    def forward(ctx, projections, angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z,
                pixel_size, volume_shape, voxel_size, volume_offset):
        dense = [t.contiguous() for t in (projections, angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z)]
        ctx.save_for_backward(*dense)
        ctx.setup = (pixel_size, voxel_size, volume_offset)
        # the calling thread's kernel settings, reused by the backward (which autograd may run
        # on another thread)
        ctx.settings = _native.CallSettings()
        return _native.bp_forward(dense[0], dense[1:], pixel_size, *volume_shape, voxel_size, *volume_offset)

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_volume):
        projections, *geometry = ctx.saved_tensors
        pixel_size, voxel_size, volume_offset = ctx.setup
        grad_projections, grad_geometry = _native.bp_backward(
            grad_volume.contiguous(), projections, geometry, pixel_size, voxel_size, *volume_offset,
            opts=ctx.settings.options(_native.TUNING_GATHER))
        wanted = ctx.needs_input_grad
        grads = [grad_projections if wanted[0] else None]
        grads += [grad_geometry[k] if wanted[1 + k] else None for k in range(8)]
        return (*grads, None, None, None, None)


def backproject(projections, angles, sid, sod, det_u0, det_v0,
                src_x, src_y, src_z, pixel_size, volume_shape, voxel_size,
                volume_offset=(0.0, 0.0, 0.0)):
    """Voxel-driven, distance-weighted backprojection of [A, U, V] projections into a
    [D, H, W] volume (z, y, x order).

    From every view, each voxel takes the bilinear detector value where the line from the
    source through the voxel meets the detector, times the FDK weight ((sod + n) / (sod + n -
    y'))^2 (n: the source offset along the central ray, y': the voxel's coordinate along it in
    the view's rotated frame).  Views whose line misses the detector interior add nothing.
    There is no filtering and no normalisation over views -- ``reconstruct_fdk`` adds those.

    ``projections`` and the eight per-view geometry tensors ([A] each: angles in radians; sid
    = source to detector, sod = source to rotation axis; det_u0 / det_v0 = the pixel indices of
    the central ray; src_x / src_y / src_z = source offsets along the world axes) are float32
    tensors on one GPU.  Lengths, ``pixel_size``, ``voxel_size`` and ``volume_offset`` (x, y, z)
    share one unit.  First-order gradients flow to the projections and the geometry.
    """
    tensors = (projections, angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z)
    _arguments.gpu_float32(zip(("projections",) + _GEOMETRY, tensors), projections)
    if projections.ndim != 3:
        raise ValueError("projections must have shape [A,U,V]")
    views, rows, cols = projections.shape
    if views < 1 or rows < 2 or cols < 2:
        raise ValueError("projections requires A >= 1 and U,V >= 2")
    _arguments.one_value_per_view(zip(_GEOMETRY, tensors[1:]), views)
    pixel = _positive_scalar(pixel_size, "pixel_size")
    voxel = _positive_scalar(voxel_size, "voxel_size")
    dims = _arguments.volume_dims(volume_shape)
    _arguments.backprojector_limits(projections.numel(), dims)
    offset = _arguments.world_offset(volume_offset)
    return _Backproject.apply(*tensors, pixel, dims, voxel, offset)

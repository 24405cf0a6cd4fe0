"""Cone-beam ray-integral projector with a first-order autograd node.

Public names, signatures, defaults and errors are those of the reference package's
``cbct_backprojector/source_only_forward_projector.py`` (``project`` and
``DifferentiableSourceOnlyForwardProjector``); the arithmetic is ``cbct_fp_forward_ex`` /
``cbct_fp_backward_ex`` of ``libcbct_hip.so``.
"""

from collections import namedtuple

from torch.autograd import Function
from torch.autograd.function import once_differentiable

from . import _arguments, _native
from .operator import _positive_scalar

_TENSOR_NAMES = ("volume", "angles", "sid", "sod", "det_u0_arr", "det_v0_arr",
                 "src_x_arr", "src_y_arr", "src_z_arr")

# The checked scalar settings of one projection.
_Setup = namedtuple("_Setup", "pixel U V step voxel offset")


def _validate_inputs(volume, angles, sid, sod, det_u0_arr, det_v0_arr,
                     src_x_arr, src_y_arr, src_z_arr, pixel_size,
                     detector_u, detector_v, step_size, voxel_size,
                     off_x, off_y, off_z):
    """Check every argument of ``project`` in the reference's order (tensors, shapes, detector
    size, ROI offsets, then the three positive scalars) and return
    (pixel_size, detector_u, detector_v, step_size, voxel_size, off_x, off_y, off_z) as plain
    numbers (interface of cbct_backprojector/source_only_forward_projector.py:30-71)."""
    tensors = (volume, angles, sid, sod, det_u0_arr, det_v0_arr, src_x_arr, src_y_arr, src_z_arr)
    _arguments.gpu_float32(zip(_TENSOR_NAMES, tensors), volume)
    if volume.ndim != 3 or min(volume.shape) < 1:
        raise ValueError("volume must have a nonempty [D,H,W] shape")
    if angles.ndim != 1 or angles.numel() < 1:
        raise ValueError("angles must have a nonempty shape [A]")
    _arguments.one_value_per_view(zip(_TENSOR_NAMES[1:], tensors[1:]), angles.numel())
    rows = _arguments.detector_extent("detector_u", detector_u)
    cols = _arguments.detector_extent("detector_v", detector_v)
    offsets = tuple(_arguments.finite_number(name, value)
                    for name, value in (("off_x", off_x), ("off_y", off_y), ("off_z", off_z)))
    return (_positive_scalar(pixel_size, "pixel_size"), rows, cols,
            _positive_scalar(step_size, "step_size"), _positive_scalar(voxel_size, "voxel_size"),
            *offsets)


class DifferentiableSourceOnlyForwardProjector(Function):
    """Autograd node of the projector.  ``apply`` takes the arguments of ``project``
    positionally; the three ROI offsets may be left out (they default to 0), as in the
    reference (source_only_forward_projector.py:74-116).  Gradients: the volume (the
    projector's exact transpose, computed only when the volume needs one) and the eight
    geometry tensors (computed only when one of them needs one)."""

    @staticmethod
    def forward(ctx, volume, angles, sid, sod, det_u0_arr, det_v0_arr,
                src_x_arr, src_y_arr, src_z_arr, pixel_size,
                detector_u, detector_v, step_size, voxel_size,
                off_x=0.0, off_y=0.0, off_z=0.0):
        checked = _validate_inputs(volume, angles, sid, sod, det_u0_arr, det_v0_arr,
                                   src_x_arr, src_y_arr, src_z_arr, pixel_size,
                                   detector_u, detector_v, step_size, voxel_size, off_x, off_y, off_z)
        setup = _Setup(*checked[:5], checked[5:])
        dense = [t.contiguous() for t in (volume, angles, sid, sod, det_u0_arr, det_v0_arr,
                                          src_x_arr, src_y_arr, src_z_arr)]
        ctx.save_for_backward(*dense)
        ctx.setup = setup
        # the calling thread's kernel settings, reused by the backward (which autograd may run
        # on another thread)
        ctx.settings = _native.CallSettings()
        # how many gradients backward must return: 14 when apply() omitted the offsets, else 17
        ctx.arity = len(ctx.needs_input_grad)
        return _native.fp_forward(dense[0], dense[1:], setup.pixel, setup.U, setup.V, setup.step, setup.voxel,
                                  *setup.offset, opts=ctx.settings.options(_native.TUNING_GATHER))

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_projection):
        volume, *geometry = ctx.saved_tensors
        setup = ctx.setup
        wanted = ctx.needs_input_grad
        need_volume, need_geometry = bool(wanted[0]), any(wanted[1:9])
        slot = _native.TUNING_SCATTER if need_volume else _native.TUNING_GATHER
        grad_volume, grad_geometry = _native.fp_backward(
            grad_projection.contiguous(), volume, geometry, setup.pixel, setup.step, setup.voxel,
            need_volume, *setup.offset, compute_grad_geometry=need_geometry, opts=ctx.settings.options(slot))
        grads = [grad_volume if need_volume else None]
        grads += [grad_geometry[k] if wanted[1 + k] else None for k in range(8)]
        grads += [None] * (ctx.arity - len(grads))
        return tuple(grads)


def project(volume, angles, sid, sod, det_u0_arr, det_v0_arr,
            src_x_arr, src_y_arr, src_z_arr, pixel_size,
            detector_u, detector_v, step_size=0.2, voxel_size=1.0,
            off_x=0.0, off_y=0.0, off_z=0.0):
    """Line integrals of ``volume`` along the rays of a cone-beam scan, as [A, U, V] projections.

    Geometry of view a: the source sits at (-sod sin b + src_x, sod cos b + src_y, src_z) with b
    = angles[a]; the detector plane stays at distance sod - sid from the rotation axis on the
    far side (src_* move the source only), and pixel (u, v) is at ((u - det_u0) pixel_size,
    (v - det_v0) pixel_size) in that plane.  ``volume`` is [D, H, W] (z, y, x); voxel (i, j, k)
    has its centre at ((k - W/2) voxel_size + off_x, (j - H/2) voxel_size + off_y,
    (i - D/2) voxel_size + off_z).  Each ray samples the trilinearly interpolated volume every
    ``step_size`` from the point where it enters the volume box, and the projection is the sum
    of the samples times ``step_size``.

    All nine tensors are float32 on one GPU (geometry tensors: shape [A]); lengths share one
    unit.  The gradients are first-order and piecewise: they ignore the jumps where a ray gains
    or loses a sample, crosses into another interpolation cell or changes its entry face.
    Scalars and ROI offsets get no gradient.  This projector's transpose is its own adjoint,
    not ``backproject``, which is distance-weighted for FDK.
    """
    return DifferentiableSourceOnlyForwardProjector.apply(
        volume, angles, sid, sod, det_u0_arr, det_v0_arr, src_x_arr, src_y_arr, src_z_arr,
        pixel_size, detector_u, detector_v, step_size, voxel_size, off_x, off_y, off_z)

"""Chunked FDK initialisation for a uniformly sampled circular scan.

Mirrors cbct_backprojector/filtering.py of the reference (same names, argument
meaning, validation and error types).  The preweight, ramp-Hann filter and the
chunked backprojection run in ``libcbct_hip.so`` (``cbct_fdk_filter``,
``cbct_fdk_reconstruct``): the filter is the reference's zero-padded FFT
convolution restated exactly as a direct linear convolution along U.

As in the reference, the cosine preweight uses centred detector coordinates
without detector/source offsets; this is the project's FDK approximation for
complete, uniformly sampled circular scans.  The filter runs on the GPU only.

Two keyword-only extensions go beyond what the reference supports
(filtering.py:1-8 and docs/USAGE.md:332 leave them to the user):
``short_scan=True`` reconstructs a scan over pi plus the fan angle with Parker
weights (angles increasing, range > pi; weight of a view = the mean angular
step instead of pi/A), and ``offset_preweight=True`` centres the cosine
preweight on each view's det_u0/det_v0 instead of the detector's middle.
"""

import math
import operator

import torch

from . import _native


def _check_projections(projections):
    if not isinstance(projections, torch.Tensor):
        raise TypeError("projections must be a torch.Tensor")
    if projections.ndim != 3 or any(size < 1 for size in projections.shape):
        raise ValueError("projections must have a nonempty [A, U, V] shape")
    if projections.dtype != torch.float32:
        raise TypeError("projections must have dtype torch.float32")


def _positive_scalar(name, value):
    value = float(value)
    if not math.isfinite(value) or value <= 0:
        raise ValueError(f"{name} must be finite and positive")
    return value


def _check_alpha(alpha):
    alpha = _positive_scalar("alpha", alpha)
    if alpha > 1:
        raise ValueError("alpha must satisfy 0 < alpha <= 1")
    return alpha


def _filter_distance(name, value, projections):
    distance = torch.as_tensor(
        value, dtype=projections.dtype, device=projections.device
    ).reshape(-1)
    if distance.numel() == 1:
        distance = distance.expand(projections.shape[0])
    elif distance.numel() != projections.shape[0]:
        raise ValueError(f"{name} must be a scalar or have one value per view")
    if not bool(torch.isfinite(distance).all()) or bool((distance <= 0).any()):
        raise ValueError(f"{name} must contain finite positive distances")
    return distance.contiguous()


def fdk_filter(projections, pixel_size, sod, sid, alpha=0.8, *, det_u0=None, det_v0=None, angles=None):
    """Apply centred cosine weighting and a Hann-windowed discrete ramp filter (filtering.py:54-104).

    Args:
        projections: Float32 line-integral tensor with layout ``[A, U, V]`` on a GPU.
        pixel_size: Square detector pixel spacing in the geometry's length unit.
        sod: Source-to-origin distance, scalar or one value per view.
        sid: Source-to-detector distance, scalar or one value per view.
        alpha: Hann cutoff as a fraction of the Nyquist frequency, in ``(0, 1]``.

    Returns:
        A contiguous float32 tensor with the same shape and device: the
        filtered projections, equal to the reference's filtering along U with
        zero padding to the next power of two at least 2U, followed by the
        ``pixel_size * sid / sod`` scale factor.

    Keyword-only (not in the reference): ``det_u0``/``det_v0`` ([A] tensors, both or
    neither) centre the preweight per view; ``angles`` ([A], increasing, range > pi)
    applies Parker short-scan weights before the ramp filter.
    """
    _check_projections(projections)
    pixel_size = _positive_scalar("pixel_size", pixel_size)
    alpha = _check_alpha(alpha)
    sod = _filter_distance("sod", sod, projections)
    sid = _filter_distance("sid", sid, projections)
    if projections.device.type != "cuda":
        raise ValueError("fdk_filter runs on the GPU only; move projections to a GPU device")
    if (det_u0 is None) != (det_v0 is None):
        raise ValueError("det_u0 and det_v0 must be given together")
    extra = []
    for name, values in (("det_u0", det_u0), ("det_v0", det_v0), ("angles", angles)):
        if values is not None:
            values = _per_view(name, values, projections)
        extra.append(values)
    if angles is not None:
        _check_short_scan(extra[2])
    return _native.fdk_filter(projections.contiguous(), pixel_size, sod, sid, alpha, *extra)


def _per_view(name, values, projections):
    values = torch.as_tensor(values, dtype=projections.dtype, device=projections.device).reshape(-1)
    if values.numel() != projections.shape[0]:
        raise ValueError(f"{name} must have one value per view")
    if not bool(torch.isfinite(values).all()):
        raise ValueError(f"{name} must contain only finite values")
    return values.contiguous()


def _check_short_scan(angles):
    if angles.numel() < 2 or not bool((angles[1:] > angles[:-1]).all()):
        raise ValueError("a short-scan reconstruction needs strictly increasing angles")
    if float(angles[-1] - angles[0]) <= math.pi:
        raise ValueError("a short-scan reconstruction needs angles spanning more than pi")


@torch.no_grad()
def reconstruct_fdk(
    projections,
    angles,
    sid,
    sod,
    det_u0,
    det_v0,
    src_x,
    src_y,
    src_z,
    pixel_size,
    volume_shape,
    voxel_size,
    volume_offset=(0.0, 0.0, 0.0),
    alpha=0.8,
    angle_chunk_size=32,
    *,
    short_scan=False,
    offset_preweight=False,
):
    """Filter and backproject in view chunks, then apply ``pi / A`` and clamp (filtering.py:107-198).

    ``projections`` is a float32 GPU tensor ``[A, U, V]``. Each geometry argument
    from ``angles`` through ``src_z`` must be a float32 tensor ``[A]`` on the same
    device; angles are radians and detector centres are pixel indices.
    ``volume_shape`` is ``(D, H, W)``; ``volume_offset`` follows ``backproject``.

    The chunks are summed exactly as the reference orders them (each chunk from
    zero, then added to the running volume), inside one fused kernel.

    Keyword-only (not in the reference): ``short_scan`` (Parker weights; the angles must
    increase and span more than pi; each view is weighted by the mean angular step) and
    ``offset_preweight`` (preweight centred on det_u0/det_v0).
    """
    _check_projections(projections)
    if not projections.is_cuda:
        raise ValueError("reconstruct_fdk requires CUDA projections")
    if min(projections.shape[1:]) < 2:
        raise ValueError("reconstruct_fdk requires detector dimensions U and V >= 2")
    num_angles = projections.shape[0]
    geometry = (angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z)
    names = ("angles", "sid", "sod", "det_u0", "det_v0", "src_x", "src_y", "src_z")
    for name, values in zip(names, geometry):
        if not isinstance(values, torch.Tensor):
            raise TypeError(f"{name} must be a torch.Tensor of shape [A]")
        if values.ndim != 1 or values.numel() != num_angles:
            raise ValueError(f"{name} must have shape [{num_angles}]")
        if values.dtype != projections.dtype or values.device != projections.device:
            raise ValueError(f"{name} must match the projections' dtype and device")
    if not bool(torch.isfinite(torch.stack(geometry)).all()):
        raise ValueError("geometry must contain only finite values")
    if bool((sid <= 0).any()) or bool((sod <= 0).any()):
        raise ValueError("sid and sod must contain positive distances")

    pixel_size = _positive_scalar("pixel_size", pixel_size)
    voxel_size = _positive_scalar("voxel_size", voxel_size)
    alpha = _check_alpha(alpha)
    try:
        angle_chunk_size = operator.index(angle_chunk_size)
        volume_shape = tuple(operator.index(size) for size in volume_shape)
    except TypeError as error:
        raise TypeError("angle_chunk_size and volume dimensions must be integers") from error
    if angle_chunk_size < 1:
        raise ValueError("angle_chunk_size must be at least 1")
    if len(volume_shape) != 3 or any(size < 1 for size in volume_shape):
        raise ValueError("volume_shape must contain three positive dimensions (D, H, W)")
    if math.prod(volume_shape) > 2**31 - 1:
        raise ValueError("volume exceeds the backprojector's 32-bit indexing limit")
    if max(volume_shape[:2]) > 8 * 65535:
        raise ValueError("D,H exceed the backprojector's 3D CUDA grid limits")
    if min(num_angles, angle_chunk_size) * math.prod(projections.shape[1:]) > 2**31 - 1:
        raise ValueError("Reduce angle_chunk_size to fit the backprojector's 32-bit indexing limit")
    volume_offset = tuple(float(value) for value in volume_offset)
    if len(volume_offset) != 3 or not all(math.isfinite(value) for value in volume_offset):
        raise ValueError("volume_offset must contain three finite values")

    if short_scan:
        _check_short_scan(angles)
    tensors = tuple(values.contiguous() for values in geometry)
    return _native.fdk_reconstruct(projections.contiguous(), tensors, pixel_size, *volume_shape,
                                   voxel_size, *volume_offset, alpha, angle_chunk_size,
                                   bool(short_scan), bool(offset_preweight))

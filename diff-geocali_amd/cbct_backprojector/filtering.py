"""FDK initialisation: cosine preweight, Hann-windowed ramp filter, chunked backprojection.

Public names, signatures, defaults and errors are those of the reference package's
``cbct_backprojector/filtering.py`` (``fdk_filter``, ``reconstruct_fdk``).  The arithmetic runs
in ``libcbct_hip.so``: the reference filters with a zero-padded FFT along U (pad = the power of
two at or above 2U), which, cropped to the first U outputs, is exactly a linear convolution with
2U - 1 taps; the library computes those taps and convolves directly (``cbct_fdk_filter``), and
``cbct_fdk_reconstruct`` fuses filter, chunked backprojection, the pi/A scale and the clamp.

The method is the reference's approximation for complete, uniformly sampled circular scans:
the cosine preweight is centred on the detector's middle and ignores detector and source
offsets.  Two keyword-only options go further (the reference leaves them out,
filtering.py:1-8, docs/USAGE.md:332): ``short_scan=True`` applies Parker weights for a scan over
more than pi (increasing angles; each view then weighs the mean angular step instead of pi/A),
and ``offset_preweight=True`` centres the preweight on each view's det_u0 / det_v0.

Device: the filter runs on the GPU only.  The reference's ``fdk_filter`` also filters CPU
tensors with torch.fft; this library has no CPU implementation of the filter (its operators
run on the GPU alone), so a CPU tensor is refused with a ValueError instead.
"""

import math
import operator

import torch

from . import _arguments, _native

_GEOMETRY = ("angles", "sid", "sod", "det_u0", "det_v0", "src_x", "src_y", "src_z")


def _check_projections(projections):
    """A non-empty [A, U, V] float32 tensor (interface of filtering.py:17-24)."""
    if not isinstance(projections, torch.Tensor):
        raise TypeError("projections must be a torch.Tensor")
    if projections.ndim != 3 or min(projections.shape) < 1:
        raise ValueError("projections must have a nonempty [A, U, V] shape")
    if projections.dtype != torch.float32:
        raise TypeError("projections must have dtype torch.float32")


def _check_alpha(alpha):
    """The Hann cut-off as a fraction of Nyquist, in (0, 1] (interface of filtering.py:34-38)."""
    value = _arguments.positive_float("alpha", alpha)
    if value > 1:
        raise ValueError("alpha must satisfy 0 < alpha <= 1")
    return value


def _filter_distance(name, value, projections):
    """sod or sid as one finite positive distance per view, on the projections' device: a
    scalar (or one-element tensor) is repeated for every view (interface of filtering.py:41-51;
    the check reads the values back to the host, as the reference does)."""
    views = projections.shape[0]
    distance = torch.as_tensor(value, dtype=projections.dtype, device=projections.device).reshape(-1)
    if distance.numel() not in (1, views):
        raise ValueError(f"{name} must be a scalar or have one value per view")
    if distance.numel() == 1:
        distance = distance.expand(views)
    if not bool(torch.isfinite(distance).all()) or bool((distance <= 0).any()):
        raise ValueError(f"{name} must contain finite positive distances")
    return distance.contiguous()


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


def _check_filter_size(projections):
    rows = projections.shape[1]
    if rows > _arguments.FDK_MAX_U:
        raise ValueError(f"U exceeds the filter's {_arguments.FDK_MAX_U}-pixel limit")
    if projections.numel() > _arguments.INT32_MAX:
        raise ValueError("projections exceed the library's 32-bit indexing limit")


def fdk_filter(projections, pixel_size, sod, sid, alpha=0.8, *, det_u0=None, det_v0=None, angles=None):
    """Weighted, ramp-filtered projections for FDK backprojection.

    Each value p[a, u, v] of the line integrals ``projections`` ([A, U, V], float32, on a GPU)
    is first weighted by sid / sqrt(sid^2 + x_u^2 + y_v^2), with (x_u, y_v) the pixel's
    position relative to the detector's middle; then every detector row is convolved along U
    with the discrete ramp kernel (1 / (4 pixel^2) at 0, -1 / (pi n pixel)^2 at odd n),
    Hann-windowed in frequency with cut-off ``alpha`` times Nyquist, exactly as the
    reference's zero-padded FFT filter does; finally the result is scaled by
    pixel_size * sid / sod.  ``sod`` and ``sid`` are scalars or one value per view.

    Keyword-only, beyond the reference: ``det_u0`` / ``det_v0`` ([A], both or neither) centre
    the preweight on each view's central-ray pixel; ``angles`` ([A], increasing, spanning more
    than pi) multiplies Parker short-scan weights in before filtering.

    Returns a new contiguous float32 tensor of the same shape on the same device.
    """
    _check_projections(projections)
    pixel = _arguments.positive_float("pixel_size", pixel_size)
    cutoff = _check_alpha(alpha)
    sod_a = _filter_distance("sod", sod, projections)
    sid_a = _filter_distance("sid", sid, projections)
    if projections.device.type != "cuda":
        raise ValueError("fdk_filter runs on the GPU only; move projections to a GPU device")
    _check_filter_size(projections)
    if (det_u0 is None) != (det_v0 is None):
        raise ValueError("det_u0 and det_v0 must be given together")
    extras = [None if v is None else _per_view(name, v, projections)
              for name, v in (("det_u0", det_u0), ("det_v0", det_v0), ("angles", angles))]
    if extras[2] is not None:
        _check_short_scan(extras[2])
    return _native.fdk_filter(projections.contiguous(), pixel, sod_a, sid_a, cutoff, *extras)


def _check_fdk_geometry(geometry, projections):
    """reconstruct_fdk's per-view geometry checks (interface of filtering.py:143-157)."""
    views = projections.shape[0]
    for name, values in zip(_GEOMETRY, geometry):
        if not isinstance(values, torch.Tensor):
            raise TypeError(f"{name} must be a torch.Tensor of shape [A]")
        if values.ndim != 1 or values.numel() != views:
            raise ValueError(f"{name} must have shape [{views}]")
        if values.dtype != projections.dtype or values.device != projections.device:
            raise ValueError(f"{name} must match the projections' dtype and device")
    if not bool(torch.isfinite(torch.stack(geometry)).all()):
        raise ValueError("geometry must contain only finite values")
    sid, sod = geometry[1], geometry[2]
    if bool((sid <= 0).any()) or bool((sod <= 0).any()):
        raise ValueError("sid and sod must contain positive distances")


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
    """FDK reconstruction of a circular scan into a [D, H, W] volume.

    Views are processed in chunks of ``angle_chunk_size``: each chunk is filtered
    (``fdk_filter``) and backprojected (``backproject``) into a volume that starts from zero
    and is then added to the running total; the total is multiplied by pi / A and negative
    values are set to zero.  The library does all of it in one call, summing in that same
    order.  Arguments are those of ``backproject`` (projections [A, U, V] float32 on a GPU,
    eight [A] geometry tensors of the same dtype and device, ``volume_shape`` (D, H, W),
    ``volume_offset`` (x, y, z)) plus the filter's ``alpha``.  Runs without autograd.

    Keyword-only, beyond the reference: ``short_scan`` (Parker weights; angles increasing and
    spanning more than pi; each view weighs the mean angular step instead of pi / A) and
    ``offset_preweight`` (preweight centred on det_u0 / det_v0).
    """
    _check_projections(projections)
    if not projections.is_cuda:
        raise ValueError("reconstruct_fdk requires CUDA projections")
    if min(projections.shape[1:]) < 2:
        raise ValueError("reconstruct_fdk requires detector dimensions U and V >= 2")
    geometry = (angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z)
    _check_fdk_geometry(geometry, projections)
    pixel = _arguments.positive_float("pixel_size", pixel_size)
    voxel = _arguments.positive_float("voxel_size", voxel_size)
    cutoff = _check_alpha(alpha)
    try:
        chunk = operator.index(angle_chunk_size)
        dims = tuple(operator.index(n) for n in volume_shape)
    except TypeError as error:
        raise TypeError("angle_chunk_size and volume dimensions must be integers") from error
    if chunk < 1:
        raise ValueError("angle_chunk_size must be at least 1")
    if len(dims) != 3 or min(dims) < 1:
        raise ValueError("volume_shape must contain three positive dimensions (D, H, W)")
    if math.prod(dims) > _arguments.INT32_MAX:
        raise ValueError("volume exceeds the backprojector's 32-bit indexing limit")
    if dims[0] > _arguments.BP_MAX_DEPTH:
        raise ValueError(f"D exceeds the backprojector's launch limit of {_arguments.BP_MAX_DEPTH} voxels")
    _check_filter_size(projections)
    offset = tuple(float(x) for x in volume_offset)
    if len(offset) != 3 or not all(math.isfinite(x) for x in offset):
        raise ValueError("volume_offset must contain three finite values")
    if short_scan:
        _check_short_scan(angles)
    dense = tuple(values.contiguous() for values in geometry)
    return _native.fdk_reconstruct(projections.contiguous(), dense, pixel, *dims, voxel, *offset, cutoff,
                                   chunk, bool(short_scan), bool(offset_preweight))

"""Argument checks shared by the public operators.

Exception types and message texts are part of the cbct_backprojector interface -- callers and
tests match on them -- so each check raises what the reference package raises for the same
input (cited per function).  The size limits are this library's own: 32-bit element offsets in
the kernels and the launch-grid dimensions of the backprojector.
"""

import math
from numbers import Integral, Real

import torch

INT32_MAX = 2**31 - 1
# The backprojection kernels cover z in workgroups of 64 voxels along the launch grid's second
# dimension (at most 65535 workgroups), so a volume may be at most this deep.
BP_MAX_DEPTH = 64 * 65535
# The FDK filter keeps a detector row's taps in registers and LDS for rows of up to 4096 pixels.
FDK_MAX_U = 4096


def positive_number(value, name):
    """float(value) of a finite positive real number that is not a bool
    (cbct_backprojector/operator.py:11-17)."""
    if isinstance(value, bool) or not isinstance(value, Real):
        raise TypeError(f"{name} must be a positive Python number")
    number = float(value)
    if not (math.isfinite(number) and number > 0):
        raise ValueError(f"{name} must be finite and positive")
    return number


def positive_float(name, value):
    """float(value) (any value float() accepts), finite and positive
    (cbct_backprojector/filtering.py:27-31; note the argument order)."""
    number = float(value)
    if not (math.isfinite(number) and number > 0):
        raise ValueError(f"{name} must be finite and positive")
    return number


def finite_number(name, value):
    """float(value) of a finite real number that is not a bool
    (cbct_backprojector/source_only_forward_projector.py:59-64)."""
    if isinstance(value, bool) or not isinstance(value, Real):
        raise TypeError(f"{name} must be a finite Python number")
    if not math.isfinite(value):
        raise ValueError(f"{name} must be finite")
    return float(value)


def detector_extent(name, size):
    """int(size) of a positive integer that is not a bool
    (cbct_backprojector/source_only_forward_projector.py:53-57)."""
    if isinstance(size, bool) or not isinstance(size, Integral):
        raise TypeError(f"{name} must be a positive integer")
    if size < 1:
        raise ValueError(f"{name} must be a positive integer")
    return int(size)


def gpu_float32(named, anchor):
    """Every (name, tensor) is a float32 tensor on the GPU that holds ``anchor``
    (cbct_backprojector/operator.py:89-95, source_only_forward_projector.py:38-44).  ``anchor``
    is the first tensor of ``named``; it is only read once it has passed its own checks."""
    for name, tensor in named:
        if not isinstance(tensor, torch.Tensor):
            raise TypeError(f"{name} must be a torch.Tensor")
        if tensor.device.type != "cuda" or tensor.dtype != torch.float32:
            raise ValueError(f"{name} must be a CUDA float32 tensor")
        if tensor.device != anchor.device:
            raise ValueError("All tensors must be on the same CUDA device")


def one_value_per_view(named, views):
    """Every (name, tensor) has shape [views] (cbct_backprojector/operator.py:101-103)."""
    for name, tensor in named:
        if tuple(tensor.shape) != (views,):
            raise ValueError(f"{name} must have shape [{views}]")


def volume_dims(volume_shape):
    """(D, H, W) of three positive integers, bools excluded
    (cbct_backprojector/operator.py:106-112)."""
    dims = tuple(volume_shape)
    if len(dims) != 3 or any(isinstance(n, bool) or not isinstance(n, Integral) or n < 1 for n in dims):
        raise ValueError("volume_shape must contain three positive integers (D,H,W)")
    return tuple(int(n) for n in dims)


def world_offset(volume_offset):
    """(x, y, z) of three finite real numbers, bools excluded
    (cbct_backprojector/operator.py:119-125)."""
    values = tuple(volume_offset)
    if len(values) != 3 or any(isinstance(x, bool) or not isinstance(x, Real) or not math.isfinite(x)
                               for x in values):
        raise ValueError("volume_offset must contain three finite numbers (x,y,z)")
    return tuple(float(x) for x in values)


def backprojector_limits(projection_elements, dims):
    """This library's backprojector limits: 32-bit element offsets into the projections and the
    volume, and the depth its launch grid covers."""
    if projection_elements > INT32_MAX or math.prod(dims) > INT32_MAX:
        raise ValueError("Input or volume exceeds the kernel's 32-bit indexing limit")
    if dims[0] > BP_MAX_DEPTH:
        raise ValueError(f"D exceeds the backprojector's launch limit of {BP_MAX_DEPTH} voxels")

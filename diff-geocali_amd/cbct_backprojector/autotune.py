"""Autotuner for the slab projector's brick shape.

The slab kernels march a 16x16 tile of rays through the volume in slabs of
``slab_thickness`` voxel planes along the principal axis; each slab's brick
(the bounding box of the tile's rays inside the slab) is staged in LDS when it
fits ``lds_floats`` floats, otherwise the slab is halved.  ``layout`` chooses
how the 64 lanes of a wave cover the tile (0: 8x8 blocks, 1: interleaved
rows).  Every setting gives identical forward projections and the same
gradients to FP32 rounding; only the speed changes, with the magnification, the
volume and detector sizes and the sampling step.

``tune_projector`` and ``tune_calibration`` time every candidate on a uniformly
spaced subset of views (work is linear in the number of views) and keep the
fastest in a JSON cache keyed by the operation and the shape class, so a shape
is tuned once per machine.  The cache path is ``$CBCT_AUTOTUNE_CACHE`` or
``~/.cache/cbct_hip/autotune.json``.
"""

import itertools
import json
import os

import torch

from . import _native, calibration

SLAB_THICKNESS = (4, 8, 12, 16)
LDS_FLOATS = (2048, 3840, 4096, 5120, 8192)
LAYOUT = (0, 1)
# The library's built-in settings per tuning slot (gathering kernels, volume-gradient adjoint).
DEFAULT_GATHER = (12, 5120, 0)
DEFAULT = DEFAULT_SCATTER = (12, 3840, 1)
_SLOT = {"forward": 0, "geometry": 0, "calibration": 0, "adjoint": 1}


def default_for(op):
    """The built-in setting of the slot that ``op`` runs in."""
    return DEFAULT_SCATTER if _SLOT[op] == 1 else DEFAULT_GATHER


def candidates():
    """Every (slab_thickness, lds_floats, layout) the tuner tries."""
    return list(itertools.product(SLAB_THICKNESS, LDS_FLOATS, LAYOUT))


# The batched calibration launch gives every workgroup the largest LDS box of all its problems, so
# one problem tuned to a large box lowers the occupancy of all of them: its tunings stay within
# the box that keeps seven workgroups (the batched kernel's register budget) on a CU.
BATCH_MAX_LDS_FLOATS = 5120


def batch_candidates():
    return [c for c in candidates() if c[1] <= BATCH_MAX_LDS_FLOATS]


def cache_path():
    return os.environ.get("CBCT_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "cbct_hip", "autotune.json"))


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key, value):
    path = cache_path()
    data = _load()
    data[key] = value
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + f".{os.getpid()}"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass                    # a read-only home only loses the cache, not the result


def shape_key(op, volume_shape, U, V, step_size, voxel_size, sid, sod):
    """Shape class: sizes rounded to multiples of 16, step in voxels and magnification to 2 digits."""
    r16 = lambda n: int(-(-int(n) // 16) * 16)
    D, H, W = volume_shape
    return (f"{op}:vol{r16(D)}x{r16(H)}x{r16(W)}:det{r16(U)}x{r16(V)}:"
            f"step{step_size / voxel_size:.2f}:mag{float(sid) / float(sod):.2f}")


def _median_ms(fn, reps):
    fn()                                        # warm-up (and first-touch of the workspace)
    times = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    times.sort()
    return times[len(times) // 2]


def _view_subset(A, views):
    n = max(1, min(int(views), A))
    return torch.linspace(0, A - 1, n).round().long().unique()


def tune_projector(volume, geometry, pixel_size, U, V, step_size, voxel_size, op="adjoint", views=16, reps=3,
                   use_cache=True, space=None):
    """Fastest (slab_thickness, lds_floats, layout) for ``op`` in {"forward", "adjoint",
    "geometry"} (adjoint with / without the volume gradient) on this shape class."""
    if op not in ("forward", "adjoint", "geometry"):
        raise ValueError("op must be 'forward', 'adjoint' or 'geometry'")
    g = calibration.stack_geometry(geometry).to(volume.device)
    key = shape_key(op, volume.shape, U, V, step_size, voxel_size, g[1].mean().item(), g[2].mean().item())
    if use_cache and space is None:
        hit = _load().get(key)
        if hit:
            return tuple(hit["best"])
    idx = _view_subset(g.shape[1], views).to(volume.device)
    sub = [x.contiguous() for x in g[:, idx]]
    n = len(idx)
    args = (pixel_size, step_size, voxel_size, 0., 0., 0.)
    if op != "forward":
        gen = torch.Generator(device=volume.device).manual_seed(3)
        up = torch.randn((n, U, V), generator=gen, device=volume.device)

    def run(opts):
        if op == "forward":
            _native.fp_forward(volume, sub, pixel_size, U, V, step_size, voxel_size, 0., 0., 0., opts=opts)
        else:
            _native.fp_backward(up, volume, sub, pixel_size, step_size, voxel_size, op == "adjoint",
                                0., 0., 0., opts=opts)

    timings = {}
    for c in (space or candidates()):
        # each candidate is passed to its calls explicitly: no setting of this or any other
        # thread changes while the tuner runs
        opts = _native.options(tuning=(1, *c), deterministic=False)
        timings[c] = _median_ms(lambda: run(opts), reps)
    best = min(timings, key=timings.get)
    if use_cache and space is None:
        _store(key, {"best": list(best), "ms_per_view": timings[best] / n,
                     "default_ms_per_view": timings.get(default_for(op), float("nan")) / n})
    return best


def tune_calibration(problem, views=8, reps=2, use_cache=True, space=None):
    """Fastest tuning of the fused calibration step for this problem's shape class; the result
    is also stored in ``problem.tuning``."""
    A, U, V = problem.measured.shape
    g = problem.geometry
    key = shape_key("calibration-batch", problem.volume.shape, U, V, problem.step_size, problem.voxel_size,
                    g[1].mean().item(), g[2].mean().item())
    hit = _load().get(key) if use_cache and space is None else None
    if hit:
        problem.tuning = tuple(hit["best"])
        return problem.tuning
    idx = _view_subset(A, views).to(problem.volume.device)
    sub = calibration.CalibrationProblem(problem.volume, problem.measured[idx].contiguous(), g[:, idx],
                                         problem.pixel_size, problem.step_size, problem.voxel_size,
                                         problem.volume_offset)
    timings = {}
    for c in (space or batch_candidates()):
        sub.tuning = c
        timings[c] = _median_ms(lambda: calibration.calibrate([sub], steps=0, use_graph=False), reps)
    best = min(timings, key=timings.get)
    if use_cache and space is None:
        _store(key, {"best": list(best), "ms_per_view": timings[best] / len(idx),
                     "default_ms_per_view": timings.get(DEFAULT_GATHER, float("nan")) / len(idx)})
    problem.tuning = best
    return best


def apply(tuning, op=None):
    """Make ``tuning`` the calling thread's projector setting of the slot that ``op`` ("forward",
    "geometry", "adjoint", "calibration") runs in, or of both slots (``op=None``)."""
    _native.set_fp_tuning(1, *tuning, op=None if op is None else _SLOT[op])


def options_for(tuning, op):
    """cbct_fp_options that run ``op`` with ``tuning`` (slab kernels), for passing to one call."""
    return _native.options(_SLOT[op], tuning=(1, *tuning))

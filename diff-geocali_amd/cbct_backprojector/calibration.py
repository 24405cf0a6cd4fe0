"""Gradient-based per-view geometry calibration on the HIP library.

The reference stops at the differentiable operators and leaves the calibration
loop to user code (examples/autograd.py:62-63; the full pipeline is excluded,
README.md:8).  This module is that loop, MI355X-native: each problem is a known
volume and a measured projection stack, and its eight per-view geometry arrays
are fitted by Adam to the mean squared reprojection residual

    L(g) = mean over the A*U*V rays of (project(volume, g) - measured)^2 .

``calibrate`` runs many problems of uneven size together: one step of all of
them is four launches (frames, one fused march that gives each ray's
projection, residual and geometry gradient, a fixed-order per-view reduction,
the Adam update), and all steps can be captured in one hipGraph
(``cbct_calib_run``).

``reprojection_error_px`` is the accuracy measure used to assess a fit: the RMS
over views and nine fixed points (the volume's corners and centre) of the
detector-position difference, in pixels, between two geometries, with the
point-to-detector map of the backprojector
(csrc/src_cbct_backprojector_cuda.cu:48-72).
"""

import ctypes
import math

import torch

from . import _native

GEOMETRY_NAMES = ("angles", "sid", "sod", "det_u0", "det_v0", "src_x", "src_y", "src_z")

# Default Adam learning rates per parameter group, in the parameters' units
# (rad, mm, mm, px, px, mm, mm, mm): about a tenth of the typical miscalibration.
DEFAULT_LR = (3e-4, 1.0, 0.6, 0.15, 0.15, 0.05, 0.05, 0.05)


class _CalibProblem(ctypes.Structure):
    _fields_ = [
        ("volume", ctypes.c_void_p), ("D", ctypes.c_int64), ("H", ctypes.c_int64), ("W", ctypes.c_int64),
        ("voxel_size", ctypes.c_float), ("off_x", ctypes.c_float), ("off_y", ctypes.c_float),
        ("off_z", ctypes.c_float),
        ("measured", ctypes.c_void_p), ("A", ctypes.c_int64), ("U", ctypes.c_int64), ("V", ctypes.c_int64),
        ("pixel_size", ctypes.c_float), ("step_size", ctypes.c_float),
        ("geometry", ctypes.c_void_p), ("adam_m", ctypes.c_void_p), ("adam_v", ctypes.c_void_p),
        ("loss_history", ctypes.c_void_p), ("grad_out", ctypes.c_void_p),
        ("slab_thickness", ctypes.c_int), ("lds_floats", ctypes.c_int), ("layout", ctypes.c_int),
    ]


def stack_geometry(geometry):
    """[8, A] float32 tensor from a dict or a sequence of eight [A] tensors."""
    if isinstance(geometry, dict):
        geometry = [geometry[name] for name in GEOMETRY_NAMES]
    if isinstance(geometry, torch.Tensor):
        if geometry.ndim != 2 or geometry.shape[0] != 8:
            raise ValueError("geometry tensor must have shape [8, A]")
        return geometry.to(torch.float32).contiguous()
    if len(geometry) != 8:
        raise ValueError("geometry needs eight per-view tensors")
    return torch.stack([torch.as_tensor(g).reshape(-1).to(torch.float32) for g in geometry]).contiguous()


class CalibrationProblem:
    """One scan to calibrate: a known volume, its measured projections and a starting geometry."""

    def __init__(self, volume, measured, geometry, pixel_size, step_size, voxel_size,
                 volume_offset=(0.0, 0.0, 0.0), tuning=None):
        if volume.dtype != torch.float32 or volume.ndim != 3 or volume.device.type != "cuda":
            raise ValueError("volume must be a [D,H,W] float32 GPU tensor")
        if measured.dtype != torch.float32 or measured.ndim != 3 or measured.device != volume.device:
            raise ValueError("measured must be an [A,U,V] float32 tensor on the volume's device")
        self.volume = volume.contiguous()
        self.measured = measured.contiguous()
        self.geometry = stack_geometry(geometry).to(volume.device).clone()
        if self.geometry.shape[1] != measured.shape[0]:
            raise ValueError("geometry must have one value per view of measured")
        self.pixel_size = float(pixel_size)
        self.step_size = float(step_size)
        self.voxel_size = float(voxel_size)
        self.volume_offset = tuple(float(x) for x in volume_offset)
        self.tuning = tuning              # (slab_thickness, lds_floats, layout) or None
        A = measured.shape[0]
        self.adam_m = torch.zeros((8, A), dtype=torch.float32, device=volume.device)
        self.adam_v = torch.zeros_like(self.adam_m)
        self.loss_history = None
        self.grad = None

    def _struct(self, nsteps, want_grad):
        self.loss_history = torch.zeros(nsteps + 1, dtype=torch.float32, device=self.volume.device)
        self.grad = torch.zeros_like(self.geometry) if want_grad else None
        D, H, W = self.volume.shape
        A, U, V = self.measured.shape
        # this problem's tuning, or the calling thread's setting for the gathering kernels: the
        # library is never left to read its process-wide default
        t = self.tuning or _native.get_fp_tuning(_native.TUNING_GATHER)[1:]
        return _CalibProblem(
            self.volume.data_ptr(), D, H, W, self.voxel_size, *self.volume_offset,
            self.measured.data_ptr(), A, U, V, self.pixel_size, self.step_size,
            self.geometry.data_ptr(), self.adam_m.data_ptr(), self.adam_v.data_ptr(),
            self.loss_history.data_ptr(), self.grad.data_ptr() if want_grad else None,
            int(t[0]), int(t[1]), int(t[2]))


def calibrate(problems, steps=20, lr=DEFAULT_LR, betas=(0.9, 0.999), eps=1e-8, use_graph=True,
              want_grad=False, lr_decay=1.0, first_step=0, solver="adam", lm_lambda=1e-3):
    """Run ``steps`` optimisation steps on every problem together; problems are updated in
    place (``problem.geometry``, ``problem.loss_history`` of length steps+1).

    solver="adam": Adam on the loss gradient with per-group learning rates ``lr``; step t
    (counted from ``first_step``; continue a fit by passing the steps already taken) uses
    ``lr * lr_decay**t``; the moments live in ``problem.adam_m/v``; with ``want_grad`` the
    gradient at the final geometry is returned in ``problem.grad``.

    solver="lm": Levenberg-Marquardt per view on the Gauss-Newton normal equations formed by
    the same fused march (initial damping ``lm_lambda``); ``loss_history`` holds the accepted
    loss after each iteration and the geometry ends at the accepted iterate."""
    if isinstance(problems, CalibrationProblem):
        problems = [problems]
    problems = list(problems)
    if not problems:
        raise ValueError("no problems")
    if solver not in ("adam", "lm"):
        raise ValueError("solver must be 'adam' or 'lm'")
    device = problems[0].volume.device
    if any(p.volume.device != device for p in problems):
        raise ValueError("all problems must be on one device")
    lib = _native.lib()
    lm = solver == "lm"
    arr = (_CalibProblem * len(problems))(*[p._struct(steps, want_grad and not lm) for p in problems])
    lr_arr = (ctypes.c_float * 8)(*[float(x) for x in lr])
    with torch.cuda.device(device):
        size_fn = lib.cbct_calib_lm_workspace_bytes if lm else lib.cbct_calib_workspace_bytes
        nbytes = size_fn(ctypes.addressof(arr), len(problems))
        ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=device)
        stream = torch.cuda.current_stream(device).cuda_stream
        if lm:
            status = lib.cbct_calib_run_lm(ctypes.addressof(arr), len(problems), int(steps), float(lm_lambda),
                                           1 if use_graph else 0, ws.data_ptr(), nbytes, stream)
        else:
            status = lib.cbct_calib_run(ctypes.addressof(arr), len(problems), int(steps), int(first_step),
                                        ctypes.addressof(lr_arr), float(lr_decay), float(betas[0]),
                                        float(betas[1]), float(eps), 1 if use_graph else 0, ws.data_ptr(), nbytes,
                                        stream)
    if status != 0:
        raise RuntimeError(lib.cbct_last_error().decode())
    return problems


# Parameters of calibrate_autograd: the detector is fitted through its physical shifts
# (delta_u, delta_v, in length units) and detector_centers, as the reference's docs suggest
# (docs/USAGE.md:304, cbct_backprojector/operator.py:32-45).
AUTOGRAD_PARAMETERS = ("angles", "sid", "sod", "delta_u", "delta_v", "src_x", "src_y", "src_z")


def detector_shifts(det_u0, det_v0, detector_u, detector_v, pixel_size):
    """(delta_u, delta_v) of detector centres: the inverse of ``detector_centers``."""
    return (((detector_u - 1) / 2.0 - det_u0) * pixel_size, ((detector_v - 1) / 2.0 - det_v0) * pixel_size)


def calibrate_autograd(volume, measured, geometry, pixel_size, step_size, voxel_size, steps=20, lr=None,
                       fit=AUTOGRAD_PARAMETERS, betas=(0.9, 0.999), eps=1e-8, volume_offset=(0.0, 0.0, 0.0)):
    """The plain calibration loop through the public, differentiable API: leaf tensors for the
    fitted parameters (the detector through its physical shifts delta_u / delta_v, mapped to
    centre indices by ``detector_centers`` inside the graph), ``project``, the mean squared
    reprojection residual, ``backward()`` and one ``torch.optim.Adam`` step per iteration, with one
    learning rate per parameter group.

    ``geometry``: the starting geometry ([8, A] or eight [A] tensors, in the projector's order with
    det_u0 / det_v0 as centre indices).  ``lr``: learning rates of AUTOGRAD_PARAMETERS (default:
    DEFAULT_LR with the detector rates converted to length units, pixel-rate x pixel_size).
    ``fit``: the parameters to fit (the others stay fixed).  Returns (geometry [8, A] as centre
    indices, {"delta_u": ..., "delta_v": ...}, loss history [steps + 1]).

    Adam divides each gradient by its own running RMS, so fitting delta_u with rate r follows the
    same trajectory as fitting det_u0 = (U-1)/2 - delta_u / pixel_size with rate r / pixel_size
    (up to the eps term): the fused batched ``calibrate`` is the fast form of this loop."""
    from .operator import detector_centers
    from .source_only_forward_projector import project

    g0 = stack_geometry(geometry).to(volume.device)
    A, U, V = measured.shape
    if lr is None:
        lr = list(DEFAULT_LR)
        lr[3] *= pixel_size
        lr[4] *= pixel_size
    du0, dv0 = detector_shifts(g0[3], g0[4], U, V, pixel_size)
    start = {"angles": g0[0], "sid": g0[1], "sod": g0[2], "delta_u": du0, "delta_v": dv0,
             "src_x": g0[5], "src_y": g0[6], "src_z": g0[7]}
    params = {k: v.detach().clone().requires_grad_(k in fit) for k, v in start.items()}
    groups = [{"params": [params[k]], "lr": float(lr[i])} for i, k in enumerate(AUTOGRAD_PARAMETERS) if k in fit]
    opt = torch.optim.Adam(groups, betas=betas, eps=eps)
    history = []

    def loss_now():
        u0, v0 = detector_centers(params["delta_u"], params["delta_v"], U, V, pixel_size)
        proj = project(volume, params["angles"], params["sid"], params["sod"], u0, v0, params["src_x"],
                       params["src_y"], params["src_z"], pixel_size, U, V, step_size, voxel_size, *volume_offset)
        return ((proj - measured) ** 2).mean()

    for _ in range(steps):
        opt.zero_grad(set_to_none=True)
        loss = loss_now()
        loss.backward()
        history.append(loss.detach())
        opt.step()
    with torch.no_grad():
        history.append(loss_now())
        u0, v0 = detector_centers(params["delta_u"], params["delta_v"], U, V, pixel_size)
        out = torch.stack([params["angles"], params["sid"], params["sod"], u0, v0, params["src_x"],
                           params["src_y"], params["src_z"]]).detach()
    shifts = {"delta_u": params["delta_u"].detach(), "delta_v": params["delta_v"].detach()}
    return out, shifts, torch.stack(history)


def fixed_points(volume_shape, voxel_size, volume_offset=(0.0, 0.0, 0.0)):
    """The volume's eight corner voxel centres and its centre, world coordinates [9, 3] (x, y, z)."""
    D, H, W = volume_shape
    pts = []
    for z in (0, D - 1):
        for y in (0, H - 1):
            for x in (0, W - 1):
                pts.append(((x - W * 0.5) * voxel_size + volume_offset[0],
                            (y - H * 0.5) * voxel_size + volume_offset[1],
                            (z - D * 0.5) * voxel_size + volume_offset[2]))
    pts.append((volume_offset[0] - 0.5 * voxel_size, volume_offset[1] - 0.5 * voxel_size,
                volume_offset[2] - 0.5 * voxel_size))
    return torch.tensor(pts, dtype=torch.float64)


def detector_positions(geometry, points, pixel_size):
    """(u_idx, v_idx) [A, n] of world points under each view's geometry, FP64
    (the map of csrc/src_cbct_backprojector_cuda.cu:48-72)."""
    g = stack_geometry(geometry).to(torch.float64)
    P = points.to(device=g.device, dtype=torch.float64)
    ang, sid, sod, u0, v0, sx, sy, sz = [x[:, None] for x in g]
    c, s = torch.cos(ang), torch.sin(ang)
    px = P[None, :, 0] * c + P[None, :, 1] * s
    py = -P[None, :, 0] * s + P[None, :, 1] * c
    pz = P[None, :, 2]
    dsu = sx * c + sy * s
    dsn = -sx * s + sy * c
    t = (sid + dsn) / (sod + dsn - py)
    u = (dsu + t * (px - dsu)) / pixel_size + u0
    v = (sz + t * (pz - sz)) / pixel_size + v0
    return u, v


def reprojection_error_px(geometry, reference, volume_shape, voxel_size, pixel_size,
                          volume_offset=(0.0, 0.0, 0.0)):
    """RMS over views and the nine fixed points of the detector-position difference (pixels)."""
    pts = fixed_points(volume_shape, voxel_size, volume_offset)
    u1, v1 = detector_positions(geometry, pts, pixel_size)
    u2, v2 = detector_positions(reference, pts, pixel_size)
    return math.sqrt(float(((u1 - u2) ** 2 + (v1 - v2) ** 2).mean()))


def synthetic_problem(spec, device, perturb=True, tuning=None):
    """A calibration problem on a seeded ellipsoid phantom: the measured projections are the
    projector's output at a perturbed geometry (returned as ``true_geometry``), and the fit
    starts from the nominal circular scan.  ``spec``: D, A, U, V, pixel, voxel, sod, sid, seed."""
    from . import phantom
    D, A, U, V = spec["D"], spec["A"], spec["U"], spec["V"]
    vs, pix = spec["voxel"], spec["pixel"]
    params = phantom.ellipsoid_parameters(spec["seed"], 12, half_width=D * vs / 2)
    volume = phantom.rasterize_ellipsoids_on(params, (D, D, D), vs, device)
    nominal = phantom.circular_geometry(A, spec["sid"], spec["sod"], U, V, pix)
    true = phantom.perturb_geometry(nominal, spec["seed"] + 1) if perturb else nominal
    true_list = [true[k].to(device) for k in GEOMETRY_NAMES]
    step = spec.get("step", vs / 2)
    measured = _native.fp_forward(volume, true_list, pix, U, V, step, vs, 0.0, 0.0, 0.0)
    problem = CalibrationProblem(volume, measured, nominal, pix, step, vs, tuning=tuning)
    problem.true_geometry = stack_geometry(true_list)
    return problem

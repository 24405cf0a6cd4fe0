"""Synthetic scans: seeded ellipsoid phantoms, circular-scan geometry and the
per-view geometry perturbations used by the benchmarks and the calibration code.

An ellipsoid is 11 numbers (cx, cy, cz, ax, ay, az, qw, qx, qy, qz, density) in
world coordinates: centre, semi-axes, a unit rotation quaternion and a
density.  A phantom is the sum of its ellipsoids' indicator functions sampled
at the voxel centres ``(i - N/2) * voxel_size + offset`` (the convention of
csrc/src_cbct_backprojector_cuda.cu:37-39).
"""

import math

import numpy as np
import torch

GEOMETRY_NAMES = ("angles", "sid", "sod", "det_u0", "det_v0", "src_x", "src_y", "src_z")


def ellipsoid_parameters(seed, count=12, half_width=1.0):
    """``count`` random ellipsoids: centres U[-0.6,0.6]*R, semi-axes U[0.05,0.4]*R,
    uniformly random rotations, densities U[0,1] (R = ``half_width``)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    centres = rng.uniform(-0.6, 0.6, (count, 3)) * half_width
    axes = rng.uniform(0.05, 0.4, (count, 3)) * half_width
    quats = rng.normal(size=(count, 4))
    quats /= np.linalg.norm(quats, axis=1, keepdims=True)
    dens = rng.uniform(0.0, 1.0, (count, 1))
    return np.concatenate([centres, axes, quats, dens], axis=1)


def quaternion_matrix(q):
    w, x, y, z = np.asarray(q, np.float64) / np.linalg.norm(q)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def rasterize_ellipsoids(params, shape, voxel_size, offset=(0.0, 0.0, 0.0)):
    """float32 [D,H,W] phantom; each ellipsoid only visits its bounding box."""
    D, H, W = shape
    vol = np.zeros((D, H, W), np.float32)
    xs = (np.arange(W) - W * 0.5) * voxel_size + offset[0]
    ys = (np.arange(H) - H * 0.5) * voxel_size + offset[1]
    zs = (np.arange(D) - D * 0.5) * voxel_size + offset[2]
    for e in np.asarray(params, np.float64).reshape(-1, 11):
        c, ax, R, rho = e[0:3], e[3:6], quaternion_matrix(e[6:10]), e[10]
        reach = np.abs(R) @ ax      # half-extent of the bounding box along each world axis
        sel = []
        for coords, k in ((xs, 0), (ys, 1), (zs, 2)):
            idx = np.nonzero(np.abs(coords - c[k]) <= reach[k])[0]
            sel.append(idx)
        if any(len(s) == 0 for s in sel):
            continue
        ix, iy, iz = sel
        Z, Y, X = np.meshgrid(zs[iz] - c[2], ys[iy] - c[1], xs[ix] - c[0], indexing="ij")
        P = np.stack([X, Y, Z], axis=-1) @ R     # local coordinates (R^T p for row vectors)
        inside = ((P / ax) ** 2).sum(-1) <= 1.0
        vol[iz[0]:iz[-1] + 1, iy[0]:iy[-1] + 1, ix[0]:ix[-1] + 1] += (rho * inside).astype(np.float32)
    return vol


def rasterize_ellipsoids_on(params, shape, voxel_size, device, offset=(0.0, 0.0, 0.0)):
    """The phantom of ``rasterize_ellipsoids`` evaluated with torch on ``device`` (FP64
    inside test), as a float32 [D,H,W] tensor; for large synthetic batches."""
    D, H, W = shape
    vol = torch.zeros((D, H, W), dtype=torch.float32, device=device)
    f64 = dict(dtype=torch.float64, device=device)
    xs = (torch.arange(W, **f64) - W * 0.5) * voxel_size + offset[0]
    ys = (torch.arange(H, **f64) - H * 0.5) * voxel_size + offset[1]
    zs = (torch.arange(D, **f64) - D * 0.5) * voxel_size + offset[2]
    for e in np.asarray(params, np.float64).reshape(-1, 11):
        c, ax, R, rho = e[0:3], e[3:6], quaternion_matrix(e[6:10]), e[10]
        reach = np.abs(R) @ ax
        sel = [torch.nonzero((coords - c[k]).abs() <= reach[k]).flatten()
               for coords, k in ((xs, 0), (ys, 1), (zs, 2))]
        if any(s.numel() == 0 for s in sel):
            continue
        ix, iy, iz = sel
        Z, Y, X = torch.meshgrid(zs[iz] - c[2], ys[iy] - c[1], xs[ix] - c[0], indexing="ij")
        P = torch.stack([X, Y, Z], dim=-1) @ torch.tensor(R, **f64)
        inside = ((P / torch.tensor(ax, **f64)) ** 2).sum(-1) <= 1.0
        z0, y0, x0 = int(iz[0]), int(iy[0]), int(ix[0])
        vol[z0:z0 + len(iz), y0:y0 + len(iy), x0:x0 + len(ix)] += (rho * inside).to(torch.float32)
    return vol


def circular_geometry(A, sid, sod, U, V, pixel_size, device="cpu", start=0.0):
    """Eight float32 [A] tensors of a uniform full-circle scan with a centred detector."""
    angles = torch.arange(A, dtype=torch.float64) * (2.0 * math.pi / A) + start
    full = lambda value: torch.full((A,), float(value), dtype=torch.float64)
    geometry = {
        "angles": angles,
        "sid": full(sid),
        "sod": full(sod),
        "det_u0": full((U - 1) / 2.0),
        "det_v0": full((V - 1) / 2.0),
        "src_x": full(0.0),
        "src_y": full(0.0),
        "src_z": full(0.0),
    }
    return {k: v.to(device=device, dtype=torch.float32) for k, v in geometry.items()}


def perturb_geometry(geometry, seed, angle_deg=0.2, center_px=1.5, src_mm=0.5, distance_rel=0.01):
    """Per-view perturbations: angle + N(0, angle_deg), det_u0/v0 + N(0, center_px),
    src x/y/z + N(0, src_mm), sid and sod * (1 + N(0, distance_rel))."""
    g = {k: v.detach().to("cpu", torch.float64).clone() for k, v in geometry.items()}
    A = g["angles"].numel()
    gen = torch.Generator().manual_seed(int(seed))
    n = lambda: torch.randn(A, generator=gen, dtype=torch.float64)
    g["angles"] += n() * math.radians(angle_deg)
    g["det_u0"] += n() * center_px
    g["det_v0"] += n() * center_px
    g["src_x"] += n() * src_mm
    g["src_y"] += n() * src_mm
    g["src_z"] += n() * src_mm
    g["sid"] *= 1.0 + n() * distance_rel
    g["sod"] *= 1.0 + n() * distance_rel
    device = next(iter(geometry.values())).device
    return {k: v.to(device=device, dtype=torch.float32) for k, v in g.items()}


def geometry_list(geometry):
    return [geometry[k] for k in GEOMETRY_NAMES]


def calibration_specs(count, seed, side=(96, 320), views=(90, 720), detector=(128, 768), voxel_size=0.5,
                      sod=600.0, sid=1100.0):
    """Sizes of ``count`` uneven calibration problems: cubic volumes of log-uniform side,
    a uniform-integer number of views, square detectors of uniform-integer side.  The pixel
    size is chosen so that the magnified volume diagonal fills 95% of the detector."""
    rng = np.random.default_rng(int(seed))
    specs = []
    for i in range(count):
        D = int(round(math.exp(rng.uniform(math.log(side[0]), math.log(side[1])))))
        A = int(rng.integers(views[0], views[1] + 1))
        U = int(rng.integers(detector[0], detector[1] + 1))
        pixel = (sid / sod) * math.sqrt(2.0) * D * voxel_size / (0.95 * U)
        specs.append({"D": D, "A": A, "U": U, "V": U, "pixel": pixel, "voxel": voxel_size, "sod": sod,
                      "sid": sid, "seed": int(seed) * 1000 + i})
    return specs

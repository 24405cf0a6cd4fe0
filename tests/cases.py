"""Shared test cases.

The two small fixtures restate the inputs of the reference's own check scripts
(examples/check_forward_operator.py:13-17,108-133 and
examples/check_operator.py:13-16,67-94): the same sizes, geometry values,
envelope volume and upstream gradients, including their non-contiguous
(transposed) layouts.  ``tests/golden/fixtures.json`` holds the same inputs as
data together with FP64 expected outputs.
"""

import json
import math
import os

import numpy as np
import torch

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures.json")
NAMES9_FP = ("volume", "angles", "sid", "sod", "det_u0", "det_v0", "src_x", "src_y", "src_z")
NAMES9_BP = ("projections", "angles", "sid", "sod", "det_u0", "det_v0", "src_x", "src_y", "src_z")


def load_golden():
    with open(GOLDEN) as fh:
        return json.load(fh)


# ---- forward-projector fixture (check_forward_operator.make_case) ----------
FP_U, FP_V = 5, 4
FP_PIXEL, FP_VOXEL, FP_STEP = 0.75, 0.8, 0.37
FP_OFFSET = (0.17, -0.12, 0.09)
FP_SELECTED_RAY = (0, 2, 1)
FP_SELECTED_WEIGHT = 1.3
FP_GEOMETRY = (
    [0.13, 1.28], [40.0, 41.0], [24.0, 24.7],
    [2.11, 1.87], [1.42, 1.69], [0.13, -0.18],
    [-0.11, 0.15], [0.19, -0.07],
)


def fp_case(dtype=torch.float64):
    """(volume [5,7,6] non-contiguous, 8 geometry tensors, upstream [2,5,4] non-contiguous)."""
    depth, height, width = 5, 7, 6
    z, x, y = torch.meshgrid(
        torch.arange(depth, dtype=dtype), torch.arange(width, dtype=dtype),
        torch.arange(height, dtype=dtype), indexing="ij")
    envelope = (x * (width - 1 - x) / (width - 1) ** 2 * y * (height - 1 - y) / (height - 1) ** 2
                * z * (depth - 1 - z) / (depth - 1) ** 2)
    volume = (64 * envelope * (1 + 0.06 * x + 0.04 * y + 0.08 * z)).transpose(1, 2)
    geometry = [torch.tensor(v, dtype=dtype) for v in FP_GEOMETRY]
    upstream = torch.linspace(0.2, 1.1, 2 * FP_U * FP_V, dtype=dtype)
    upstream = upstream.reshape(2, FP_V, FP_U).transpose(1, 2)
    return volume, geometry, upstream


# ---- backprojector fixture (check_operator.make_case) -----------------------
BP_PIXEL, BP_VOXEL = 0.9, 1.0
BP_SHAPE = (3, 4, 5)
BP_OFFSET = (0.27, -0.31, 0.23)
BP_GEOMETRY = (
    [0.17, 1.23, 2.51], [18.0, 18.7, 17.8], [12.0, 12.4, 11.8],
    [4.17, 3.83, 4.11], [3.13, 2.91, 3.19], [0.13, -0.21, 0.16],
    [-0.17, 0.14, 0.09], [0.11, -0.08, 0.17],
)


def bp_case(dtype=torch.float64):
    """(projections [3,9,7] non-contiguous, 8 geometry tensors, upstream [3,4,5] non-contiguous)."""
    views, du, dv = 3, 9, 7
    a = torch.arange(views, dtype=dtype)[:, None, None]
    v = torch.arange(dv, dtype=dtype)[None, :, None]
    u = torch.arange(du, dtype=dtype)[None, None, :]
    proj = (0.4 + 0.07 * a + 0.08 * u + 0.05 * v + 0.013 * u * v).transpose(1, 2)
    geometry = [torch.tensor(g, dtype=dtype) for g in BP_GEOMETRY]
    D, H, W = BP_SHAPE
    upstream = torch.linspace(0.2, 1.1, D * H * W, dtype=dtype).reshape(D, W, H).transpose(1, 2)
    return proj, geometry, upstream


# ---- plain operators at the CPU-runnable reference size ---------------------
C1 = dict(shape=(16, 20, 24), A=24, U=48, V=40, pixel=1.0, sid=180.0, sod=120.0, step=0.5, voxel=1.0)
C1_SEED = 20260101


def c1_volume():
    rng = np.random.Generator(np.random.PCG64(C1_SEED))
    return rng.random(C1["shape"], dtype=np.float32)


def c1_geometry():
    A, U, V = C1["A"], C1["U"], C1["V"]
    angles = (np.arange(A, dtype=np.float32) * np.float32(2 * math.pi / A)).astype(np.float32)
    g = [angles, np.full(A, C1["sid"], np.float32), np.full(A, C1["sod"], np.float32),
         np.full(A, (U - 1) / 2.0, np.float32), np.full(A, (V - 1) / 2.0, np.float32),
         np.zeros(A, np.float32), np.zeros(A, np.float32), np.zeros(A, np.float32)]
    return g


def to_np(t):
    return t.detach().cpu().numpy()


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))

"""Detector preprocessing and layout conversion (GPU) against the FP64 oracle."""

import numpy as np
import pytest
import torch

import oracle
from cbct_backprojector import formats

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


@pytest.mark.parametrize("dtype", ["uint16", "float32"])
@pytest.mark.parametrize("layout", ["AVU", "AUV"])
@pytest.mark.parametrize("shape", [(67, 130), (66, 260), (68, 264)])
def test_preprocess_matches_oracle(dtype, layout, shape):
    """(67, 130): odd rows, the element-wise transpose; (66, 260): even sizes, the paired
    kernel; (68, 264): rows a multiple of 4 and columns of 8, the multi-view kernel (11 views:
    one workgroup's batch of 8 views is cut short); all ragged against the 64-pixel tiles."""
    rng = np.random.default_rng(5)
    A, (R, C) = (11 if shape == (68, 264) else 3), shape
    dark = rng.uniform(80, 120, (R, C)).astype(np.float32)
    flat = rng.uniform(3000, 60000, (R, C)).astype(np.float32)
    p = rng.uniform(0.0, 5.0, (A, R, C))
    raw = dark + (flat - dark) * np.exp(-p)
    raw = np.clip(np.round(raw), 0, 65535).astype(np.uint16) if dtype == "uint16" else raw.astype(np.float32)
    raw[0, 0, :5] = 0                       # below dark
    flat[1, 3] = dark[1, 3]                 # dead pixel
    t_raw = torch.from_numpy(raw).to(DEV)
    got = formats.preprocess(t_raw, torch.tensor(dark, device=DEV), torch.tensor(flat, device=DEV), layout,
                             min_transmission=1e-5)
    ref = oracle.preprocess(raw, dark, flat, layout, min_transmission=1e-5)
    g = got.double().cpu().numpy()
    assert g.shape == ref.shape
    assert np.abs(g - ref).max() <= 2e-6 * np.abs(ref).max() + 1e-6


@pytest.mark.parametrize("shape", [(70, 129), (70, 258), (68, 136)])
def test_transpose_only_is_exact(shape):
    rng = np.random.default_rng(6)
    x = rng.standard_normal((4,) + shape).astype(np.float32)
    got = formats.to_auv(torch.tensor(x, device=DEV))
    assert np.array_equal(got.cpu().numpy(), x.transpose(0, 2, 1))


@pytest.mark.parametrize("dtype", ["uint16", "float32"])
def test_multiview_kernel_without_fields(dtype):
    """No dark / flat (0 and 1): -log(max(raw, t_min)) on the multi-view kernel's path."""
    rng = np.random.default_rng(8)
    A, R, C = 9, 64, 128
    raw = rng.uniform(0.0, 1.2, (A, R, C))
    raw = raw.astype(np.float32) if dtype == "float32" else np.round(raw * 3).astype(np.uint16)
    got = formats.preprocess(torch.from_numpy(raw).to(DEV), None, None, "AVU", min_transmission=1e-3)
    ref = oracle.preprocess(raw, None, None, "AVU", min_transmission=1e-3)
    assert np.abs(got.double().cpu().numpy() - ref).max() <= 2e-6 * np.abs(ref).max() + 1e-6

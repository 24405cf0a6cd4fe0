"""Host-side logic that needs no GPU: the uneven calibration problem sizes, the
reprojection-error measure, and the calibration entry point's argument checks."""

import ctypes
import math

import torch

from cbct_backprojector import _native, calibration, phantom


def test_calibration_specs_are_uneven_and_in_range():
    specs = phantom.calibration_specs(64, seed=4)
    sides = [s["D"] for s in specs]
    assert min(sides) >= 96 and max(sides) <= 320 and len(set(sides)) > 20
    assert all(90 <= s["A"] <= 720 and 128 <= s["U"] == s["V"] <= 768 for s in specs)
    # log-uniform side: about half of the problems below the geometric mean of the range
    below = sum(s["D"] < math.sqrt(96 * 320) for s in specs)
    assert 20 <= below <= 44
    assert phantom.calibration_specs(64, seed=4) == specs


def test_reprojection_error_zero_for_same_geometry_and_exact_for_detector_shift():
    geom = phantom.circular_geometry(36, 1100.0, 600.0, 128, 128, 1.0)
    shape = (64, 64, 64)
    assert calibration.reprojection_error_px(geom, geom, shape, 0.5, 1.0) == 0.0
    shifted = dict(geom)
    shifted["det_u0"] = geom["det_u0"] + 0.3
    shifted["det_v0"] = geom["det_v0"] - 0.4
    # a detector-centre shift moves every point by exactly (0.3, -0.4) px
    err = calibration.reprojection_error_px(shifted, geom, shape, 0.5, 1.0)
    assert abs(err - 0.5) < 1e-6


def test_detector_positions_follow_the_backprojector_map():
    """The centre of rotation projects to the detector centre; a point on the source side
    at (0, y) for angle 0 is magnified by sid / (sod - y)."""
    geom = phantom.circular_geometry(4, 1000.0, 500.0, 100, 80, 0.5)
    pts = torch.tensor([[0.0, 0.0, 0.0], [10.0, 100.0, 4.0]], dtype=torch.float64)
    u, v = calibration.detector_positions(geom, pts, 0.5)
    assert abs(u[0, 0].item() - 49.5) < 1e-9 and abs(v[0, 0].item() - 39.5) < 1e-9
    mag = 1000.0 / (500.0 - 100.0)
    assert abs(u[0, 1].item() - (49.5 + 10.0 * mag / 0.5)) < 1e-9
    assert abs(v[0, 1].item() - (39.5 + 4.0 * mag / 0.5)) < 1e-9


def test_calib_entry_point_rejects_bad_arguments_without_a_gpu():
    lib = _native.lib()
    lr = (ctypes.c_float * 8)(*([0.1] * 8))
    assert lib.cbct_calib_workspace_bytes(None, 0) == 0
    st = lib.cbct_calib_run(None, 0, 5, 0, ctypes.addressof(lr), 1.0, 0.9, 0.999, 1e-8, 0, None, 0, None)
    assert st != 0 and b"at least one problem" in lib.cbct_last_error()
    prob = calibration._CalibProblem()
    prob.D = prob.H = prob.W = 8
    prob.A = prob.U = prob.V = 4
    st = lib.cbct_calib_run(ctypes.addressof(prob), 1, -1, 0, ctypes.addressof(lr), 1.0, 0.9, 0.999, 1e-8, 0,
                            None, 0, None)
    assert st != 0 and b"nsteps" in lib.cbct_last_error()
    st = lib.cbct_calib_run(ctypes.addressof(prob), 1, 3, 0, ctypes.addressof(lr), 1.0, 1.5, 0.999, 1e-8, 0,
                            None, 0, None)
    assert st != 0 and b"Adam" in lib.cbct_last_error()
    nbytes = lib.cbct_calib_workspace_bytes(ctypes.addressof(prob), 1)
    assert nbytes > 0
    st = lib.cbct_calib_run(ctypes.addressof(prob), 1, 3, 0, ctypes.addressof(lr), 1.5, 0.9, 0.999, 1e-8, 0,
                            None, 0, None)
    assert st != 0 and b"lr_decay" in lib.cbct_last_error()
    st = lib.cbct_calib_run(ctypes.addressof(prob), 1, 3, 0, ctypes.addressof(lr), 1.0, 0.9, 0.999, 1e-8, 0,
                            None, 0, None)
    assert st != 0 and b"workspace" in lib.cbct_last_error()


def test_geometry_from_table_follows_the_csv_convention(tmp_path):
    from cbct_backprojector import formats
    from cbct_backprojector.operator import detector_centers
    A, U, V, pix = 4, 100, 80, 0.5
    rows = ["SAD_Rs,DD,delta_U,delta_V,delta_src_x,delta_src_y,delta_src_z,delta_Beta"]
    for k in range(A):
        rows.append(f"{600 + k},{1100 - k},{0.1 * k},{-0.2 * k},{0.01 * k},{0.02 * k},{0.03 * k},{0.001 * k}")
    path = tmp_path / "geom.csv"
    path.write_text("\n".join(rows) + "\n")
    g = formats.geometry_from_table(str(path), U, V, pix)
    k = torch.arange(A, dtype=torch.float64)
    assert torch.allclose(g["sod"].double(), 600 + k) and torch.allclose(g["sid"].double(), 1100 - k)
    u0, v0 = detector_centers(0.1 * k, -0.2 * k, U, V, pix)
    assert torch.allclose(g["det_u0"].double(), u0.double(), atol=1e-5)
    assert torch.allclose(g["det_v0"].double(), v0.double(), atol=1e-5)
    assert torch.allclose(g["angles"].double(), 2 * math.pi * k / A + 0.001 * k, atol=1e-6)
    assert torch.allclose(g["src_z"].double(), 0.03 * k, atol=1e-7)
    assert list(g) == ["angles", "sid", "sod", "det_u0", "det_v0", "src_x", "src_y", "src_z"]


def test_geometry_from_table_rejects_missing_or_ragged_columns():
    import pytest
    from cbct_backprojector import formats
    with pytest.raises(ValueError, match="SAD_Rs"):
        formats.geometry_from_table({"DD": [1.0]}, 4, 4, 1.0)
    with pytest.raises(ValueError, match="one value per view"):
        formats.geometry_from_table({"SAD_Rs": [1.0, 2.0], "DD": [3.0]}, 4, 4, 1.0)


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_autotune_cache_round_trip_and_shape_classes(tmp_path, monkeypatch):
    from cbct_backprojector import autotune
    monkeypatch.setenv("CBCT_AUTOTUNE_CACHE", str(tmp_path / "sub" / "cache.json"))
    k1 = autotune.shape_key("adjoint", (256, 256, 256), 768, 640, 0.25, 0.5, 1100.0, 600.0)
    k2 = autotune.shape_key("adjoint", (250, 256, 256), 760, 640, 0.25, 0.5, 1100.0, 600.0)
    k3 = autotune.shape_key("forward", (256, 256, 256), 768, 640, 0.25, 0.5, 1100.0, 600.0)
    assert k1 == k2 and k1 != k3                       # sizes round to multiples of 16
    autotune._store(k1, {"best": [12, 4096, 1], "ms_per_view": 0.3})
    autotune._store(k3, {"best": [8, 8192, 0], "ms_per_view": 0.2})
    data = autotune._load()
    assert data[k1]["best"] == [12, 4096, 1] and data[k3]["best"] == [8, 8192, 0]
    assert len(autotune.candidates()) == 40 and autotune.DEFAULT in autotune.candidates()
    assert autotune.DEFAULT_GATHER in autotune.candidates()
    batch = autotune.batch_candidates()
    assert batch and set(batch) <= set(autotune.candidates())
    assert max(c[1] for c in batch) == autotune.BATCH_MAX_LDS_FLOATS


def test_detector_shifts_invert_detector_centers():
    import torch
    from cbct_backprojector import calibration, detector_centers
    du = torch.tensor([0.3, -1.2, 0.0], dtype=torch.float64)
    dv = torch.tensor([-0.7, 0.25, 2.0], dtype=torch.float64)
    u0, v0 = detector_centers(du, dv, 48, 40, 0.8)
    su, sv = calibration.detector_shifts(u0, v0, 48, 40, 0.8)
    assert torch.allclose(su, du, atol=1e-12) and torch.allclose(sv, dv, atol=1e-12)

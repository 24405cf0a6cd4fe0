"""Every ray march ends: views that no FP32 march can follow are degenerate (zero projection,
zero gradients), and a step below the FP32 resolution of the ray parameter is refused on the
host.  CPU only -- the oracle restates the library's view guard with the same float
expressions, and runs here in a subprocess under a time limit, so a march that failed to end
would fail the test instead of hanging it.  The C ABI's refusal is a host check made before
any device call, so it is tested here too.  None of these inputs is ever run on a GPU."""

import json
import os
import subprocess
import sys

import pytest

import cases  # noqa: F401  (sets sys.path)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# A small scan with four good views and one view per failure class: a NaN angle, an infinite
# sid, a source so far away (sod = 1e9 mm) that the step is below the FP32 spacing of t
# there, and a detector centre whose corner rays overflow.
SCRIPT = r"""
import json, math, sys
import numpy as np
sys.path.insert(0, {root!r})
import oracle

rng = np.random.default_rng(5)
D, H, W, U, V = 10, 12, 14, 16, 12
vol = rng.random((D, H, W)).astype(np.float32)
A = 8
ang = (np.arange(A) * (2 * math.pi / A)).astype(np.float32)
geo = [ang, np.full(A, 90.0, np.float32), np.full(A, 60.0, np.float32),
       np.full(A, (U - 1) / 2, np.float32), np.full(A, (V - 1) / 2, np.float32),
       np.full(A, 0.1, np.float32), np.full(A, -0.2, np.float32), np.full(A, 0.05, np.float32)]
bad = {{1: (0, float("nan")), 3: (1, float("inf")), 5: (2, 1e9), 6: (3, 3e37)}}
for a, (k, x) in bad.items():
    geo[k][a] = np.float32(x)
good = [a for a in range(A) if a not in bad]
sub = [np.ascontiguousarray(g[good]) for g in geo]
out = {{}}
for mode in ("f32", "f64"):
    p = oracle.fp_forward(vol, geo, 1.0, U, V, 0.5, 1.0, mode=mode)
    ps = oracle.fp_forward(vol, sub, 1.0, U, V, 0.5, 1.0, mode=mode)
    up = rng.standard_normal((A, U, V)).astype(np.float32)
    gv, gg = oracle.fp_backward(up, vol, geo, 1.0, 0.5, 1.0, mode=mode)
    gvs, ggs = oracle.fp_backward(np.ascontiguousarray(up[good]), vol, sub, 1.0, 0.5, 1.0, mode=mode)
    S, hits = oracle.fp_count_samples((D, H, W), geo, 1.0, U, V, 0.5, 1.0, mode=mode)
    Ss, hs = oracle.fp_count_samples((D, H, W), sub, 1.0, U, V, 0.5, 1.0, mode=mode)
    out[mode] = dict(
        bad_zero=bool(all(np.all(p[a] == 0) for a in bad)),
        good_equal=bool(np.array_equal(p[good], ps)),
        good_nonzero=bool(np.all(np.abs(ps).sum(axis=(1, 2)) > 0)),
        grad_finite=bool(np.isfinite(gv).all() and np.isfinite(gg).all()),
        grad_bad_zero=bool(all(np.all(gg[:, a] == 0) for a in bad)),
        gvol_equal=float(np.abs(gv - gvs).max() / max(np.abs(gvs).max(), 1e-30)),
        ggeo_equal=float(np.abs(gg[:, good] - ggs).max() / max(np.abs(ggs).max(), 1e-30)),
        count_equal=bool(S == Ss and hits == hs and S > 0),
    )
try:
    oracle.fp_forward(vol, geo, 1.0, U, V, 1e-9, 1.0)
    out["tiny_step"] = "accepted"
except ValueError as e:
    out["tiny_step"] = str(e)
print(json.dumps(out))
"""


def test_oracle_marches_end_on_degenerate_views():
    """NaN / infinite geometry, a source too far for the step, and corner rays that overflow
    give degenerate views (zeros) in both oracle instantiations; the other views are
    unchanged, and a step of 1e-9 is refused.  A march that did not end would hit the time
    limit."""
    proc = subprocess.run([sys.executable, "-c", SCRIPT.format(root=ROOT)], capture_output=True, text=True,
                          timeout=120)
    assert proc.returncode == 0, proc.stderr
    res = json.loads(proc.stdout.strip().splitlines()[-1])
    for mode in ("f32", "f64"):
        r = res[mode]
        assert r["bad_zero"] and r["good_equal"] and r["good_nonzero"], (mode, r)
        assert r["grad_finite"] and r["grad_bad_zero"] and r["count_equal"], (mode, r)
        # the gradients of the good views are the same up to the regrouping of the per-thread
        # partial sums (FP rounding only)
        assert r["gvol_equal"] < 1e-5 and r["ggeo_equal"] < 1e-5, (mode, r)
    assert "FP32 resolution" in res["tiny_step"]


def test_oracle_guard_thresholds():
    """The guard's step bound at its edge: the smallest accepted step is 2^-22 * Tb, Tb = 2 (|O - C|
    + R) + 1; just below it the view is degenerate, at it the view is marched (and ends)."""
    code = r"""
import json, math, sys
import numpy as np
sys.path.insert(0, {root!r})
import oracle
D = H = W = 4
vol = np.ones((D, H, W), np.float32)
U = V = 4
sod = np.float32(3000.0)
geo = [np.zeros(1, np.float32), np.full(1, 4000.0, np.float32), np.full(1, sod, np.float32),
       np.full(1, 1.5, np.float32), np.full(1, 1.5, np.float32)] + [np.zeros(1, np.float32)] * 3
# source O = (0, sod, 0); box centre C = (-0.5, -0.5, -0.5); R = 0.5 * sqrt(48)
O = np.array([0.0, float(sod), 0.0]); C = np.full(3, float(np.float32(-0.5)))
R = float(np.float32(0.5 * math.sqrt(48.0)))
Tb = 2.0 * (float(np.linalg.norm(O - C)) + R) + 1.0
lo = float(np.nextafter(np.float32(Tb * 2.0 ** -22), np.float32(0)))
hi = float(np.nextafter(np.float32(Tb * 2.0 ** -22), np.float32(1)))
res = {{}}
for name, step in (("below", lo), ("above", hi)):
    S, hits = oracle.fp_count_samples((D, H, W), geo, 0.001, U, V, step, 1.0)
    res[name] = [S, hits]
print(json.dumps(res))
"""
    proc = subprocess.run([sys.executable, "-c", code.format(root=ROOT)], capture_output=True, text=True,
                          timeout=300)
    assert proc.returncode == 0, proc.stderr
    res = json.loads(proc.stdout.strip().splitlines()[-1])
    assert res["below"] == [0, 0]
    S, hits = res["above"]
    assert hits == 16 and S > 0


def test_library_refuses_a_step_below_fp32_resolution():
    """The C ABI refuses step_size = 1e-9 on a 64^3 volume (and its calibration entry point
    does for a problem), before touching any pointer or device."""
    from cbct_backprojector import _native
    L = _native.lib()
    nul = [None] * 8
    st = L.cbct_fp_forward(None, 64, 64, 64, *nul, 3, 1.0, 16, 16, 1e-9, 1.0, 0.0, 0.0, 0.0, None, None, 0, None)
    assert st == 1 and b"FP32 resolution" in L.cbct_last_error()
    st = L.cbct_fp_backward(None, None, 64, 64, 64, *nul, 3, 16, 16, 1.0, 1e-9, 1.0, 1, 0.0, 0.0, 0.0,
                            None, None, None, 0, None)
    assert st == 1 and b"FP32 resolution" in L.cbct_last_error()
    st = L.cbct_fp_count_samples(64, 64, 64, *nul, 3, 1.0, 16, 16, 1e-9, 1.0, 0.0, 0.0, 0.0, None, None, 0, None)
    assert st == 1 and b"FP32 resolution" in L.cbct_last_error()
    # the smallest accepted step at 64^3, voxel 1: 2^-22 * (64 sqrt(3) + 1)
    st = L.cbct_fp_forward(None, 64, 64, 64, *nul, 3, 1.0, 16, 16, 2.7e-5, 1.0, 0.0, 0.0, 0.0, None, None, 0, None)
    assert st == 1 and b"null" in L.cbct_last_error()


@pytest.mark.parametrize("step", [1e-9])
def test_calibration_refuses_a_step_below_fp32_resolution(step):
    from cbct_backprojector import _native, calibration
    L = _native.lib()
    p = calibration._CalibProblem()
    p.volume = p.measured = p.geometry = p.adam_m = p.adam_v = p.loss_history = 1   # never dereferenced
    p.D = p.H = p.W = 64
    p.voxel_size = 1.0
    p.A, p.U, p.V = 3, 16, 16
    p.pixel_size, p.step_size = 1.0, step
    import ctypes
    lr = (ctypes.c_float * 8)(*([0.1] * 8))
    st = L.cbct_calib_run(ctypes.byref(p), 1, 1, 0, ctypes.cast(lr, ctypes.c_void_p), 1.0, 0.9, 0.999, 1e-8, 0,
                          ctypes.c_void_p(1), 1 << 40, None)
    assert st == 1 and b"FP32 resolution" in L.cbct_last_error()

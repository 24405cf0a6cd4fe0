"""The bench line's roofline object is computed from the per-launch counters kept in bench.py's
MEASURED table (the figures of profiles/fp_slab_geometry-gradients-256c-720v.md) by the formulas
stated in bench.py's docstring: these tests recompute them by hand."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402

WL = "geometry-gradients-256c-720v"


def test_valu_peak_is_two_wave_instructions_per_cu_cycle():
    # 2 wave64 instructions x 64 lanes x 256 CUs x 2.4 GHz
    assert bench.VALU_PEAK_TOPS == pytest.approx(78.6432, rel=1e-12)


@pytest.mark.parametrize("kernel,seconds", [("adjoint", 0.1546), ("forward", 0.0606)])
def test_roofline_frac_and_shares_from_the_counters(kernel, seconds):
    m = bench.MEASURED[(WL, kernel)]
    r = bench.roofline(WL, kernel, seconds, tap_bytes=1e12)
    assert r["bound"] == "valu"
    achieved = m["valu"] * 64 / seconds / 1e12
    assert r["achieved"] == pytest.approx(achieved)
    assert r["frac"] == pytest.approx(achieved / 78.6432)
    assert 0.0 < r["frac"] < 1.0
    cu_cycles = 256 * m["grbm"] / 8
    assert r["lds_bank_conflict_share"] == pytest.approx(m["lds_conflict"] / cu_cycles)
    assert r["lds_active_share"] == pytest.approx(m["lds_active"] / cu_cycles)
    assert r["lds_issue_wait_share"] == pytest.approx(m["lds_wait"] / m["wave_qcycles"])
    # HBM bytes: FETCH_SIZE x 2 + WRITE_SIZE x 1 (KiB), the factors of a 1 GiB copy
    traffic = (m["fetch_kib"] * 2.0 + m["write_kib"]) * 1024
    assert r["traffic"] == pytest.approx(traffic)
    assert r["hbm_frac"] == pytest.approx(traffic / seconds / 1e9 / 8000.0)
    assert r["tap_model_ratio"] == pytest.approx(1e12 / seconds / 1e9 / 8000.0)


def test_profile_figures():
    """The values the profile quotes: adjoint 0.55 of VALU issue at its 154.6 ms kernel-trace
    average, LDS busy 78% and 41% in bank conflicts; forward 0.66 at 60.6 ms, LDS busy 71%."""
    adj = bench.roofline(WL, "adjoint", 0.154639, tap_bytes=None)
    fwd = bench.roofline(WL, "forward", 0.060550, tap_bytes=None)
    assert round(adj["frac"], 2) == 0.55 and round(fwd["frac"], 2) == 0.66
    assert round(adj["lds_active_share"], 2) == 0.78 and round(adj["lds_bank_conflict_share"], 2) == 0.41
    assert round(fwd["lds_active_share"], 2) == 0.71


def test_without_counters_no_roof_is_claimed():
    r = bench.roofline(WL, "adjoint", 0.16, tap_bytes=1e12, measured_ok=False)
    assert r["bound"] is None and r["frac"] is None and r["traffic"] is None

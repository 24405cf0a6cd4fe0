#!/usr/bin/env python3
"""Forward + adjoint throughput of the cone-beam operators on one MI355X.

Workload (default): a 256^3 fp32 phantom of 12 random ellipsoids, A=720 views,
768x640 detector, pixel 0.6 mm, SOD 600 mm, SID 1100 mm, voxel 0.5 mm, step
0.25 mm, per-view geometry perturbed (angle N(0,0.2 deg), centres N(0,1.5 px),
source N(0,0.5 mm), distances x(1+N(0,0.01))).  One step = one forward
projection plus one adjoint (volume gradient and the eight per-view geometry
gradients) for a fixed upstream gradient, inputs resident in HBM.

Prints ONE JSON line:
  value      GUPS = 2*D*H*W*A / (t_fwd + t_adj) / 1e9
  roofline   the roof the dominant kernel is measured against: VALU issue.  achieved =
             SQ_INSTS_VALU per launch (a constant of the workload and kernel, from rocprofv3
             --pmc, kept below) x 64 lanes / the kernel's HIP-event duration measured here;
             peak = 2 wave instructions per CU-cycle x 64 lanes x 256 CUs x 2.4 GHz = 78.6
             TOP/s (the 157.3 TFLOP/s FP32 vector peak counted per lane-instruction).  Beside
             it: the measured HBM traffic and its fraction of 8 TB/s; the LDS pipe's busy
             share, its bank-conflict share and the share of wave time spent waiting to issue
             an LDS instruction (the adjoint keeps the LDS busier than the VALU: both are
             reported); and the tap model's bytes (forward: 32 B per trilinear sample + 4 B
             per ray; adjoint: 64 B per sample + 4 B per ray + 4 B per voxel) over the
             duration, as a ratio to 8 TB/s (above 1 because LDS serves most taps: it is not
             a roof).
  cpu_baseline  the FP32 CPU oracle (OpenMP) on a bounded subset of views.
  calibration   the metric's second half (default workload, unless --no-calibration): seconds
             to 0.1 px reprojection error on 256^3 x 180-view scans from the learned warm start
             and from the plain start, for Adam and for Levenberg-Marquardt, with the network
             trained in the run (about 45 s; not counted in the times).
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "diff-geocali_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "fwd+adjoint GUPS at 256³×720 views; calib s to 0.1 px, learned vs plain"
HBM_PEAK_GBPS = 8000.0
# MI355X VALU issue peak in lane-instructions per second: 2 wave64 instructions per CU-cycle x 64
# lanes x 256 CUs x 2.4 GHz (= the published 157.3 TFLOP/s FP32 vector peak / 2 FLOP per FMA).
VALU_PEAK_TOPS = 2 * 64 * 256 * 2.4e9 / 1e12

WORKLOADS = {
    "geometry-gradients-256c-720v": dict(D=256, A=720, U=768, V=640, pixel=0.6, sod=600.0, sid=1100.0,
                                         voxel=0.5, step=0.25, perturb=True),
    "fdk-roundtrip-256c-360v": dict(D=256, A=360, U=512, V=384, pixel=0.8, sod=500.0, sid=1000.0,
                                    voxel=0.5, step=0.25, perturb=False, fdk=True),
}
CALIB_WORKLOAD = "batched-calibration-64p"
WORKLOADS[CALIB_WORKLOAD] = dict(problems=64, adam_steps=20, side=(96, 320), views=(90, 720), detector=(128, 768),
                                 voxel=0.5, sod=600.0, sid=1100.0, seed=4)
WARM_WORKLOAD = "calibration-time-256c-180v"
# one Adam schedule for both starts: calibration.DEFAULT_LR, constant (multiplier, decay per step)
WORKLOADS[WARM_WORKLOAD] = dict(train_steps=1500, pool=24, batch=2, scans=(7002, 7003, 7004), threshold=0.1,
                                passes_adam=3, passes_lm=1, adam_budget=2000, lm_budget=20, lm_lambda=1e-3,
                                adam=(1.0, 1.0))
BP_WORKLOAD = "backprojection-256c-360v"
WORKLOADS[BP_WORKLOAD] = dict(D=256, A=360, U=512, V=384, pixel=0.8, sod=500.0, sid=1000.0, voxel=0.5, step=0.25,
                              perturb=True)
PREP_WORKLOAD = "preprocess-720v-768x640-u16"
WORKLOADS[PREP_WORKLOAD] = dict(A=720, U=768, V=640)
DEFAULT_WORKLOAD = "geometry-gradients-256c-720v"
# Counters per launch of the projector kernels, from rocprofv3 --pmc runs of `bench.py --pmc-probe`
# (summaries and commands in profiles/): SQ_INSTS_VALU (wave instructions), SQ_LDS_BANK_CONFLICT and
# SQ_LDS_IDX_ACTIVE (LDS cycles stalled by bank conflicts / busy, summed over CUs), SQ_WAIT_INST_LDS
# and SQ_WAVE_CYCLES (wave quad-cycles waiting to issue an LDS instruction / resident), GRBM_GUI_ACTIVE
# (cycles, summed over the 8 XCDs), FETCH_SIZE and WRITE_SIZE (KiB).  A 1 GiB copy in the same runs
# reads 524,306 KiB of FETCH_SIZE and 1,048,576 KiB of WRITE_SIZE, so HBM bytes = FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0 (KiB x 1024).  The counts depend only on the workload, the kernel and its tuning
# (SQ_WAIT_INST_LDS and SQ_WAVE_CYCLES vary by about 1% between runs).
MEASURED = {
    ("geometry-gradients-256c-720v", "adjoint"): dict(valu=104_597_247_185, lds_conflict=38_899_405_936,
                                                      lds_active=73_355_711_832, lds_wait=56_032_938_653,
                                                      wave_qcycles=438_637_946_358, grbm=2_935_886_323,
                                                      fetch_kib=101_631_588, write_kib=137_481_984),
    ("geometry-gradients-256c-720v", "forward"): dict(valu=48_996_622_309, lds_conflict=9_670_306_160,
                                                      lds_active=26_231_495_702, lds_wait=14_547_534_868,
                                                      wave_qcycles=241_074_440_990, grbm=1_157_686_989,
                                                      fetch_kib=91_054_548, write_kib=1_470_551),
    # one fused Adam march of the 64-problem batched calibration (all problems, one launch)
    (CALIB_WORKLOAD, "march"): dict(valu=1_166_453_661_623, lds_conflict=179_186_528_075, grbm=28_067_277_858,
                                    fetch_kib=3_145_644_934, write_kib=98_684_285),
    # forward at 360 views, 512x384 (FDK round trip; HBM counters only); backprojection backward /
    # forward at 256^3 x 360 views
    ("fdk-roundtrip-256c-360v", "forward"): dict(fetch_kib=47_285_859, write_kib=296_987),
    ("backprojection-256c-360v", "bp_backward"): dict(valu=8_252_738_194, lds_conflict=896_532_168,
                                                      lds_active=3_079_634_000, lds_wait=817_399_235,
                                                      wave_qcycles=27_536_107_527, grbm=258_441_513,
                                                      fetch_kib=1_714_566, write_kib=7_138_081),
    ("backprojection-256c-360v", "bp_forward"): dict(valu=2_672_033_792, lds_conflict=0, lds_active=0, lds_wait=0,
                                                     wave_qcycles=25_513_154_256, grbm=106_890_065,
                                                     fetch_kib=5_650_494, write_kib=168_070),
}
CUS, XCDS = 256, 8


def traffic_bytes(m):
    """Measured HBM bytes per launch (None without the two counters)."""
    if m is None or "fetch_kib" not in m:
        return None
    return m["fetch_kib"] * 1024 * 2.0 + m["write_kib"] * 1024 * 1.0


def roofline(workload, kernel, seconds, tap_bytes, measured_ok=True, copy_gbps=None):
    """The roofline object of one kernel of a workload (see the module docstring)."""
    m = MEASURED.get((workload, kernel)) if measured_ok else None
    traffic = traffic_bytes(m)
    tap_gbps = tap_bytes / seconds / 1e9 if tap_bytes else None
    out = {"kernel": kernel, "traffic": traffic,
           "hbm_GBps": None if traffic is None else traffic / seconds / 1e9,
           "hbm_frac": None if traffic is None else traffic / seconds / 1e9 / HBM_PEAK_GBPS,
           "tap_model_GBps": tap_gbps, "tap_model_ratio": None if tap_gbps is None else tap_gbps / HBM_PEAK_GBPS,
           "copy_GBps": copy_gbps}
    if m is not None and "valu" in m:
        achieved = m["valu"] * 64 / seconds / 1e12
        cu_cycles = CUS * m["grbm"] / XCDS
        out.update(bound="valu", achieved=achieved, peak=VALU_PEAK_TOPS, unit="TOP/s (VALU lane-instructions)",
                   frac=achieved / VALU_PEAK_TOPS, valu_insts_per_launch=m["valu"],
                   lds_bank_conflict_share=m["lds_conflict"] / cu_cycles)
        if "lds_active" in m:
            # the LDS pipe's busy share of the launch and the share of resident wave time spent waiting
            # to issue an LDS instruction
            out.update(lds_active_share=m["lds_active"] / cu_cycles,
                       lds_issue_wait_share=m["lds_wait"] / m["wave_qcycles"])
    elif traffic is not None:
        achieved = traffic / seconds / 1e9
        out.update(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBPS, unit="GB/s", frac=achieved / HBM_PEAK_GBPS)
    else:
        out.update(bound=None, achieved=None, peak=None, unit=None, frac=None)
    return out


def cpu_identity():
    """Model name of the host CPU (from /proc/cpuinfo) and the hardware threads visible."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


PHANTOM_SEED = 20261203
PERTURB_SEED = 7


def parse():
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS))
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU-baseline sample")
    p.add_argument("--pmc-probe", action="store_true",
                   help="run a 1 GiB reference copy, one forward, one adjoint and one adjoint without geometry "
                        "gradients (for rocprofv3 --pmc runs) and exit")
    p.add_argument("--adam-steps", type=int, default=None,
                   help="Adam steps per problem of the batched-calibration workload (default 20)")
    p.add_argument("--solver", default="adam", choices=("adam", "lm"),
                   help="optimiser of the batched-calibration workload (the workload is defined with Adam)")
    p.add_argument("--train-steps", type=int, default=None,
                   help="training steps of the warm-start network in the calibration-time workload")
    p.add_argument("--deterministic", action="store_true",
                   help="bitwise-reproducible volume / projection gradients (int64 fixed-point accumulation)")
    p.add_argument("--scans", type=int, default=None,
                   help="held-out scans of the calibration-time workload (default 3)")
    p.add_argument("--no-calibration", action="store_true",
                   help="leave the calibration-time object out of the default workload's line")
    p.add_argument("--fp-tuning", default=None,
                   help="projector kernels as use_slab,slab_thickness,lds_floats,layout (default: library's)")
    return p.parse_args()


def make_inputs(w, device):
    import numpy as np
    import torch

    from cbct_backprojector import phantom
    D = w["D"]
    half = D * w["voxel"] / 2
    params = phantom.ellipsoid_parameters(PHANTOM_SEED, 12, half_width=half)
    vol_np = phantom.rasterize_ellipsoids(params, (D, D, D), w["voxel"])
    geom = phantom.circular_geometry(w["A"], w["sid"], w["sod"], w["U"], w["V"], w["pixel"])
    if w["perturb"]:
        geom = phantom.perturb_geometry(geom, PERTURB_SEED)
    glist = [geom[k].contiguous() for k in phantom.GEOMETRY_NAMES]
    vol = torch.tensor(vol_np, device=device)
    gdev = [g.to(device) for g in glist]
    gen = torch.Generator(device=device).manual_seed(11)
    upstream = torch.randn((w["A"], w["U"], w["V"]), generator=gen, device=device, dtype=torch.float32)
    return vol_np, [g.numpy() for g in glist], vol, gdev, upstream


def copy_rate_gbps(torch, device):
    n = 1 << 28   # 1 GiB of float32
    a = torch.empty(n, dtype=torch.float32, device=device).fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / 5 / 1e3
    del a, b
    return 2 * n * 4 / t / 1e9


def sample_views(A, n=256):
    """Up to n distinct views for a CPU-baseline sample, in an order whose every prefix is spread
    over the scan (bit-reversed fractions of the orbit), so stopping at the time budget still
    covers the whole orbit."""
    bits = max(1, (max(A, 2) - 1).bit_length())
    out, seen = [], set()
    for i in range(1 << bits):
        a = int(int(format(i, f"0{bits}b")[::-1], 2) * A >> bits)
        if a not in seen:
            seen.add(a)
            out.append(a)
    return out[:min(A, n)]


def cpu_baseline(w, vol_np, geom_np, budget_s):
    """FP32 CPU oracle forward + adjoint on uniformly spaced views until the budget is used."""
    import numpy as np

    import oracle
    A = w["A"]
    order = sample_views(A)
    threads = oracle.max_threads()
    rng = np.random.default_rng(11)
    done, t_total = 0, 0.0
    for a in order:
        g1 = [np.ascontiguousarray(g[a:a + 1]) for g in geom_np]
        up = rng.standard_normal((1, w["U"], w["V"])).astype(np.float32)
        t0 = time.perf_counter()
        oracle.fp_forward(vol_np, g1, w["pixel"], w["U"], w["V"], w["step"], w["voxel"], mode="f32",
                          threads=threads)
        oracle.fp_backward(up, vol_np, g1, w["pixel"], w["step"], w["voxel"], mode="f32", threads=threads)
        t_total += time.perf_counter() - t0
        done += 1
        if t_total >= budget_s:
            break
    D = w["D"]
    gups = 2.0 * D * D * D * done / t_total / 1e9
    return {"value": gups, "unit": "GUPS", "cores": threads, "kind": "port", "cpu_model": cpu_identity()[0],
            "threads_visible": cpu_identity()[1],
            "sample": f"{done} of {A} views (uniformly spaced), forward + adjoint with volume gradient, "
                      f"FP32 reference-schedule C oracle, {t_total:.1f} s"}


def cpu_baseline_fdk(w, vol_np, geom_np, budget_s):
    """FP32 CPU oracle forward projection + FDK (numpy FP64 filter, C FP32 chunked
    backprojection) of single views, uniformly spaced, until the budget is used."""
    import numpy as np

    import oracle
    A, D = w["A"], w["D"]
    order = sample_views(A)
    threads = oracle.max_threads()
    done, t_total = 0, 0.0
    for a in order:
        g1 = [np.ascontiguousarray(g[a:a + 1]) for g in geom_np]
        t0 = time.perf_counter()
        p1 = oracle.fp_forward(vol_np, g1, w["pixel"], w["U"], w["V"], w["step"], w["voxel"], mode="f32",
                               threads=threads)
        f1 = oracle.fdk_filter(p1.astype(np.float64), w["pixel"], g1[2], g1[1], 0.8).astype(np.float32)
        oracle.bp_forward(f1, g1, w["pixel"], (D, D, D), w["voxel"], mode="f32", threads=threads)
        t_total += time.perf_counter() - t0
        done += 1
        if t_total >= budget_s:
            break
    return {"value": 2.0 * D ** 3 * done / t_total / 1e9, "unit": "GUPS", "cores": threads, "kind": "port", "cpu_model": cpu_identity()[0],
            "threads_visible": cpu_identity()[1],
            "sample": f"{done} of {A} views (uniformly spaced): forward projection (FP32 C oracle), FDK filter "
                      f"(numpy FP64) and backprojection (FP32 C oracle), {t_total:.1f} s"}


def cpu_baseline_bp(w, geom_np, proj, gvol_np, budget_s):
    """FP32 CPU oracle backprojection + backward of single views, uniformly spaced."""
    import numpy as np

    import oracle
    A, D = w["A"], w["D"]
    order = sample_views(A)
    threads = oracle.max_threads()
    done, t_total = 0, 0.0
    for a in order:
        g1 = [np.ascontiguousarray(g[a:a + 1]) for g in geom_np]
        p1 = proj[a:a + 1].cpu().numpy()
        t0 = time.perf_counter()
        oracle.bp_forward(p1, g1, w["pixel"], (D, D, D), w["voxel"], mode="f32", threads=threads)
        oracle.bp_backward(gvol_np, p1, g1, w["pixel"], w["voxel"], mode="f32", threads=threads)
        t_total += time.perf_counter() - t0
        done += 1
        if t_total >= budget_s:
            break
    return {"value": 2.0 * D ** 3 * done / t_total / 1e9, "unit": "GUPS", "cores": threads, "kind": "port", "cpu_model": cpu_identity()[0],
            "threads_visible": cpu_identity()[1],
            "sample": f"{done} of {A} views (uniformly spaced): backprojection + backward, FP32 C oracle, "
                      f"{t_total:.1f} s"}


def calibration_bench(args, torch, device):
    """Batched calibration: 64 uneven problems (cubic phantoms of log-uniform side 96-320,
    90-720 views, square detectors 128-768, geometry perturbed as in the 256^3 x 720 workload),
    each fitted by 20 Adam steps in one hipGraph of persistent launches; brick shapes from the
    autotuner.  One step of this bench = one whole batched calibration (20 Adam steps plus the
    final loss evaluation = 21 fused marches per problem).  GUPS counts D^3*A voxel-view updates
    per march."""
    from cbct_backprojector import _native, autotune, calibration, phantom
    w = WORKLOADS[CALIB_WORKLOAD]
    specs = phantom.calibration_specs(w["problems"], w["seed"], w["side"], w["views"], w["detector"], w["voxel"],
                                      w["sod"], w["sid"])
    t0 = time.perf_counter()
    probs = [calibration.synthetic_problem(s, device) for s in specs]
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    for p in probs:
        autotune.tune_calibration(p, views=4, reps=1)
    tune_s = time.perf_counter() - t0
    print(f"# {len(probs)} problems set up in {setup_s:.1f} s, tuned in {tune_s:.1f} s", file=sys.stderr, flush=True)
    starts = [p.geometry.clone() for p in probs]
    samples = 0
    for p, s in zip(probs, specs):
        samples += _native.fp_count_samples((s["D"],) * 3, list(p.geometry), s["pixel"], s["U"], s["V"],
                                            p.step_size, p.voxel_size)
    nsteps = w["adam_steps"] if args.adam_steps is None else args.adam_steps

    def reset():
        for p, g in zip(probs, starts):
            p.geometry.copy_(g)
            p.adam_m.zero_()
            p.adam_v.zero_()

    for _ in range(args.warmup):
        reset()
        t0 = time.perf_counter()
        calibration.calibrate(probs, steps=nsteps, use_graph=True, solver=args.solver)
        print(f"# warm-up calibration {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    times = []
    t_wall = 0.0
    for _ in range(args.steps):
        reset()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record()
        calibration.calibrate(probs, steps=nsteps, use_graph=True, solver=args.solver)
        e1.record()
        torch.cuda.synchronize()
        t_wall += time.perf_counter() - t0
        times.append(e0.elapsed_time(e1) / 1e3)
    t = sum(times) / len(times)
    marches = nsteps + 1
    updates = sum(s["D"] ** 3 * s["A"] for s in specs) * marches
    rays = sum(s["A"] * s["U"] * s["V"] for s in specs)
    # fused march taps: 32 B per trilinear sample (the gradient reuses them) + 4 B measured per ray
    bytes_all = (32.0 * samples + 4.0 * rays) * marches
    err0 = [calibration.reprojection_error_px(g, p.true_geometry, tuple(p.volume.shape), p.voxel_size,
                                              p.pixel_size) for g, p in zip(starts, probs)]
    err1 = [calibration.reprojection_error_px(p.geometry, p.true_geometry, tuple(p.volume.shape), p.voxel_size,
                                              p.pixel_size) for p in probs]
    loss_ratio = [p.loss_history[-1].item() / p.loss_history[0].item() for p in probs]
    med = lambda x: sorted(x)[len(x) // 2]
    out = {
        "metric": METRIC, "value": updates / t / 1e9, "unit": "GUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": CALIB_WORKLOAD, "problems": len(probs), "solver": args.solver, "steps_per_problem": nsteps,
                   "s_per_batched_calibration": t, "samples_per_march": samples, "rays": rays,
                   "voxel_view_updates_per_march": updates // marches,
                   "setup_s": setup_s, "autotune_s": tune_s,
                   "reprojection_px_median_before": med(err0), "reprojection_px_median_after": med(err1),
                   "loss_ratio_median": med(loss_ratio),
                   "tunings": sorted({str(p.tuning) for p in probs})},
        # the whole graph is timed; one march is its time over the number of marches
        "roofline": roofline(CALIB_WORKLOAD, "march", t / marches, bytes_all / marches,
                             measured_ok=args.solver == "adam" and nsteps == w["adam_steps"]),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_calibration(probs, args.cpu_seconds)
    print(json.dumps(out), flush=True)


def cpu_baseline_calibration(probs, budget_s):
    """One calibration march restated on the CPU: the FP32 C oracle's forward projection and
    geometry-only adjoint (residual as upstream weight) of single views of the problems, in
    turn, until the budget is used."""
    import numpy as np

    import oracle
    threads = oracle.max_threads()
    done_updates, t_total, views = 0, 0.0, 0
    for k, p in enumerate(probs):
        A, U, V = p.measured.shape
        a = (7 * k) % A
        g1 = [np.ascontiguousarray(p.geometry[i, a:a + 1].cpu().numpy()) for i in range(8)]
        vol = p.volume.cpu().numpy()
        meas = p.measured[a:a + 1].cpu().numpy()
        t0 = time.perf_counter()
        proj = oracle.fp_forward(vol, g1, p.pixel_size, U, V, p.step_size, p.voxel_size, mode="f32",
                                 threads=threads)
        up = (2.0 / (A * U * V)) * (proj - meas)
        oracle.fp_backward(up.astype(np.float32), vol, g1, p.pixel_size, p.step_size, p.voxel_size,
                           need_volume=False, mode="f32", threads=threads)
        t_total += time.perf_counter() - t0
        done_updates += vol.size
        views += 1
        if t_total >= budget_s:
            break
    return {"value": done_updates / t_total / 1e9, "unit": "GUPS", "cores": threads, "kind": "port", "cpu_model": cpu_identity()[0],
            "threads_visible": cpu_identity()[1],
            "sample": f"one view of each of {views} problems: forward projection + geometry-only adjoint "
                      f"(FP32 C oracle), {t_total:.1f} s"}


def calibration_time(args, torch, device):
    """Seconds to 0.1 px reprojection geometry error on 256^3 x 180-view scans (detector 512^2,
    pixel 0.8 mm, geometry perturbed as in the 256^3 x 720 workload): learned warm start against
    the plain (nominal) start, for each solver.  The network is trained here first on
    projector-generated scans (training time reported, not counted).  For each held-out scan and
    method the iteration count to the threshold is found first, untimed (Adam in 10-step
    chunks); then exactly that many iterations run again from the same start inside one timed
    call, the network passes included for the learned methods.  Both starts run the same Adam
    schedule (the default per-group learning rates, constant: decaying schedules stall short of
    the threshold on some scans) and the same Levenberg-Marquardt damping; the network is
    applied three times before Adam and once before Levenberg-Marquardt, whose own iterations
    cost about as much as a network pass.  Returns the calibration object."""
    from cbct_backprojector import calibration as C
    from cbct_backprojector import warmstart as ws
    w = WORKLOADS[WARM_WORKLOAD]
    steps_train = w["train_steps"] if args.train_steps is None else args.train_steps
    fam = ws.ScanFamily()
    log = lambda m: print(f"# {m}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    data = ws.TrainingData(fam, pool=w["pool"], seed=1000, device=device)
    net = ws.WarmStartNet(seed=0, device=device)
    hist = ws.train(net, data, steps_train, P=w["batch"], lr=1e-3, lr_final=1e-5, log_every=500, log=log)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    del data
    thr = w["threshold"]
    sched = w["adam"]
    err = lambda p: C.reprojection_error_px(p.geometry, p.true_geometry, tuple(p.volume.shape), p.voxel_size,
                                            p.pixel_size)

    def reset(p, g):
        p.geometry.copy_(g)
        p.adam_m.zero_()
        p.adam_v.zero_()

    def adam_lr(sched):
        return [x * sched[0] for x in C.DEFAULT_LR]

    def adam_iters(p, g, sched):
        reset(p, g)
        chunk = 10
        for t in range(0, w["adam_budget"], chunk):
            C.calibrate([p], steps=chunk, lr=adam_lr(sched), lr_decay=sched[1], first_step=t, use_graph=False)
            if err(p) < thr:
                return t + chunk
        return None

    def lm_iters(p, g):
        for t in range(1, w["lm_budget"] + 1):
            reset(p, g)
            C.calibrate([p], steps=t, solver="lm", lm_lambda=w["lm_lambda"])
            if err(p) < thr:
                return t
        return None

    def timed(p, g_nom, n, solver, sched, warm):
        reset(p, g_nom)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if warm:
            ws.warm_start(p, net, fam, iterations=warm)
        if n:
            if solver == "lm":
                C.calibrate([p], steps=n, solver="lm", lm_lambda=w["lm_lambda"])
            else:
                C.calibrate([p], steps=n, lr=adam_lr(sched), lr_decay=sched[1])
        torch.cuda.synchronize()
        return time.perf_counter() - t0, err(p)

    rows = []
    for seed in w["scans"][:args.scans]:
        p = C.synthetic_problem(fam.spec(seed), device)
        g_nom = p.geometry.clone()
        row = {"scan": seed, "px_nominal": err(p), "px_by_network_passes": []}
        g_warm = {}
        for k in range(1, max(w["passes_adam"], w["passes_lm"]) + 1):
            reset(p, g_nom)
            ws.warm_start(p, net, fam, iterations=k)
            row["px_by_network_passes"].append(err(p))
            g_warm[k] = p.geometry.clone()
        methods = [("plain_adam", g_nom, "adam", 0), ("learned_adam", g_warm[w["passes_adam"]], "adam", w["passes_adam"]),
                   ("plain_lm", g_nom, "lm", 0), ("learned_lm", g_warm[w["passes_lm"]], "lm", w["passes_lm"])]
        for name, start, solver, warm in methods:
            n = adam_iters(p, start, sched) if solver == "adam" else lm_iters(p, start)
            t, e_t = timed(p, g_nom, n, solver, sched, warm) if n else (None, None)
            row[name] = {"iterations": n, "s": t, "px": e_t}
            log(f"scan {seed} {name}: {n} iterations, {t} s, {e_t} px")
        rows.append(row)

    def med(key):
        v = [r[key]["s"] for r in rows]
        return None if not v or any(x is None for x in v) else sorted(v)[len(v) // 2]

    ratio = lambda a, b: None if a is None or b is None else a / b
    s = {k: med(k) for k in ("plain_adam", "learned_adam", "plain_lm", "learned_lm")}

    def fastest(start):
        done = [(v, k) for k, v in s.items() if k.startswith(start) and v is not None]
        return min(done)[1] if done else None

    return {"workload": WARM_WORKLOAD, "unit": "s to 0.1 px (median over held-out scans)", "threshold_px": thr,
            "volume": [fam.D] * 3, "views": fam.A, "detector": [fam.U, fam.V], "pixel_mm": fam.pixel,
            "median_s": s,
            "speedup_same_solver": {"adam": ratio(s["plain_adam"], s["learned_adam"]),
                                    "lm": ratio(s["plain_lm"], s["learned_lm"])},
            "fastest": {start: fastest(start) for start in ("plain", "learned")},
            "network": {"parameters": net.num_parameters,
                        "passes": {"adam": w["passes_adam"], "lm": w["passes_lm"]}, "train_steps": steps_train,
                        "train_s": train_s, "train_loss_first_last": [sum(hist[:50]) / 50, sum(hist[-50:]) / 50],
                        "dtype": "bf16 GEMMs, f32 accumulation"},
            "adam_schedule": {"lr": [x * sched[0] for x in C.DEFAULT_LR], "decay_per_step": sched[1],
                              "budget_steps": w["adam_budget"]},
            "lm_lambda": w["lm_lambda"], "scans": rows}


def warm_start_bench(args, torch, device):
    """The calibration-time workload alone: its object as the JSON line's config."""
    c = calibration_time(args, torch, device)
    out = {
        "metric": METRIC, "value": c["median_s"]["learned_lm"], "unit": c["unit"], "n_gpus": 1,
        "steps": len(c["scans"]), "warmup": 0, "ms_per_step": None, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 projector / bf16 network", "data": "synthetic", "config": c,
    }
    print(json.dumps(out), flush=True)


def backprojection_bench(args, torch, device):
    """Weighted backprojection (R6) and its backward (R7) at 256^3 x 360 views, 512x384 detector,
    perturbed geometry.  One step = one backproject + one backward (projection gradient and the
    eight geometry gradients) for a fixed upstream volume gradient.  Algorithmic bytes per
    voxel-view: forward 16 (four taps); backward 16 (taps) + 16 (four scattered adds)."""
    from cbct_backprojector import _native
    if args.deterministic:
        _native.set_deterministic(True)
    w = WORKLOADS[BP_WORKLOAD]
    _, geom_np, vol, geom, _ = make_inputs(w, device)
    D, A, U, V = w["D"], w["A"], w["U"], w["V"]
    scal = (w["pixel"], U, V, w["step"], w["voxel"], 0.0, 0.0, 0.0)
    proj = _native.fp_forward(vol, geom, *scal)
    gen = torch.Generator(device=device).manual_seed(12)
    gvol = torch.randn((D, D, D), generator=gen, device=device)

    def step(ev=None):
        if ev:
            ev[0].record()
        out = _native.bp_forward(proj, geom, w["pixel"], D, D, D, w["voxel"], 0.0, 0.0, 0.0)
        if ev:
            ev[1].record()
        back = _native.bp_backward(gvol, proj, geom, w["pixel"], w["voxel"], 0.0, 0.0, 0.0)
        if ev:
            ev[2].record()
        return out, back

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_bp = sum(e[0].elapsed_time(e[1]) for e in evs) / args.steps / 1e3
    t_bb = sum(e[1].elapsed_time(e[2]) for e in evs) / args.steps / 1e3
    N = D ** 3 * A
    b_bp, b_bb = 16.0 * N + 4.0 * D ** 3, 32.0 * N + 4.0 * D ** 3
    dom = ("bp_backward", b_bb, t_bb) if t_bb >= t_bp else ("bp_forward", b_bp, t_bp)
    out = {
        "metric": METRIC, "value": 2.0 * N / (t_bp + t_bb) / 1e9, "unit": "GUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": BP_WORKLOAD, "volume": [D] * 3, "views": A, "detector": [U, V],
                   "t_bp_ms": t_bp * 1e3, "t_bp_backward_ms": t_bb * 1e3, "deterministic": args.deterministic},
        "roofline": roofline(BP_WORKLOAD, dom[0], dom[2], dom[1], measured_ok=not args.deterministic),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_bp(w, geom_np, proj, gvol.cpu().numpy(), args.cpu_seconds)
    print(json.dumps(out), flush=True)


def preprocess_bench(args, torch, device):
    """Raw uint16 frames [720, 640, 768] ([A,V,U]) + dark/flat -> line integrals [720, 768, 640]
    fp32 (dark/flat correction, -log, transpose) in one kernel; HBM-bound.  Algorithmic bytes:
    2 B read + 4 B written per pixel, plus the two fp32 fields once."""
    from cbct_backprojector import formats
    w = WORKLOADS[PREP_WORKLOAD]
    A, U, V = w["A"], w["U"], w["V"]
    gen = torch.Generator(device=device).manual_seed(3)
    raw = torch.randint(200, 60000, (A, V, U), generator=gen, device=device, dtype=torch.int32).to(torch.uint16)
    dark = torch.full((V, U), 100.0, device=device)
    flat = torch.full((V, U), 60000.0, device=device)
    for _ in range(args.warmup):
        formats.preprocess(raw, dark, flat)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        out = formats.preprocess(raw, dark, flat)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t = e0.elapsed_time(e1) / args.steps / 1e3
    nbytes = A * U * V * 6 + 2 * U * V * 4
    achieved = nbytes / t / 1e9
    del out
    print(json.dumps({
        "metric": METRIC, "value": A * U * V / t / 1e9, "unit": "Gpixel/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u16->f32", "data": "synthetic",
        "config": {"workload": PREP_WORKLOAD, "views": A, "detector": [U, V], "layout_in": "AVU",
                   "ms_per_pass": t * 1e3},
        # algorithmic bytes (6 per pixel): the kernel reads and writes each byte once
        "roofline": {"bound": "hbm", "kernel": "preprocess_views_kernel<uint16_t,8>", "achieved": achieved,
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": None},
    }), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the operators run on a single device; multi-GPU sharding is design-only).", file=sys.stderr)
        sys.exit(2)
    import torch

    from cbct_backprojector import _native
    from cbct_backprojector.source_only_forward_projector import project  # noqa: F401

    device = torch.device("cuda:0")
    torch.cuda.set_device(device)
    if args.workload == CALIB_WORKLOAD:
        return calibration_bench(args, torch, device)
    if args.workload == WARM_WORKLOAD:
        return warm_start_bench(args, torch, device)
    if args.workload == PREP_WORKLOAD:
        return preprocess_bench(args, torch, device)
    if args.workload == BP_WORKLOAD:
        return backprojection_bench(args, torch, device)
    if args.fp_tuning:
        _native.set_fp_tuning(*[int(x) for x in args.fp_tuning.split(",")])
    if args.deterministic:
        _native.set_deterministic(True)
    w = WORKLOADS[args.workload]
    vol_np, geom_np, vol, geom, upstream = make_inputs(w, device)
    D, A, U, V = w["D"], w["A"], w["U"], w["V"]
    scal = (w["pixel"], U, V, w["step"], w["voxel"], 0.0, 0.0, 0.0)
    if args.pmc_probe:
        # reference copy of known size (1 GiB read + 1 GiB written) to scale FETCH_SIZE / WRITE_SIZE
        a = torch.ones(1 << 28, device=device)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        _native.fp_forward(vol, geom, *scal)
        _native.fp_backward(upstream, vol, geom, w["pixel"], w["step"], w["voxel"], True, 0.0, 0.0, 0.0)
        # the adjoint without geometry gradients (volume scatter only), for the bound analysis
        _native.fp_backward(upstream, vol, geom, w["pixel"], w["step"], w["voxel"], True, 0.0, 0.0, 0.0,
                            compute_grad_geometry=False)
        torch.cuda.synchronize()
        print("pmc probe done", flush=True)
        return
    S = _native.fp_count_samples((D, D, D), geom, w["pixel"], U, V, w["step"], w["voxel"])

    fdk = w.get("fdk", False)
    proj_fixed = _native.fp_forward(vol, geom, *scal) if fdk else None

    def step(events=None):
        if events is not None:
            events[0].record()
        proj = _native.fp_forward(vol, geom, *scal)
        if events is not None:
            events[1].record()
        if fdk:   # reconstruct_fdk of the projections (filter + chunked backprojection)
            out = _native.fdk_reconstruct(proj_fixed, geom, w["pixel"], D, D, D, w["voxel"], 0.0, 0.0, 0.0,
                                          0.8, 32)
        else:     # adjoint: volume gradient + 8 geometry gradients
            out = _native.fp_backward(upstream, vol, geom, w["pixel"], w["step"], w["voxel"], True,
                                      0.0, 0.0, 0.0)
        if events is not None:
            events[2].record()
        return proj, out

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_fwd = sum(e[0].elapsed_time(e[1]) for e in evs) / args.steps / 1e3
    t_adj = sum(e[1].elapsed_time(e[2]) for e in evs) / args.steps / 1e3
    N = D * D * D * A
    gups = 2.0 * N / (t_fwd + t_adj) / 1e9
    rays = A * U * V
    b_fwd = 32.0 * S + 4.0 * rays
    if fdk:   # backprojection taps: 16 B per voxel-view, one 4 B volume write per chunk
        b_adj = 16.0 * D * D * D * A + 4.0 * D * D * D * math.ceil(A / 32) + 8.0 * rays
        second = "fdk_reconstruct"
    else:
        b_adj = 64.0 * S + 4.0 * rays + 4.0 * D * D * D
        second = "adjoint"
    dominant = (second, b_adj, t_adj) if t_adj >= t_fwd else ("forward", b_fwd, t_fwd)
    try:
        copy_bw = copy_rate_gbps(torch, device)
    except Exception:  # pragma: no cover
        copy_bw = None
    out = {
        "metric": METRIC,
        "value": gups,
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "volume": [D, D, D], "views": A, "detector": [U, V],
                   "pixel_mm": w["pixel"], "sod_mm": w["sod"], "sid_mm": w["sid"], "voxel_mm": w["voxel"],
                   "step_mm": w["step"], "geometry_perturbed": w["perturb"], "samples": S,
                   "t_fwd_ms": t_fwd * 1e3, ("t_fdk_ms" if fdk else "t_adj_ms"): t_adj * 1e3,
                   "fp_tuning": {"gather": list(_native.get_fp_tuning(_native.TUNING_GATHER)),
                                 "adjoint": list(_native.get_fp_tuning(_native.TUNING_SCATTER))},
                   "deterministic": args.deterministic},
        "roofline": roofline(args.workload, dominant[0], dominant[2], dominant[1],
                             measured_ok=not (args.fp_tuning or args.deterministic), copy_gbps=copy_bw),
        # the other kernel of the step, against the same roofs
        "roofline_other": (roofline(args.workload, "forward", t_fwd, b_fwd,
                                    measured_ok=not (args.fp_tuning or args.deterministic))
                           if dominant[0] != "forward" else None),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = (cpu_baseline_fdk(w, vol_np, geom_np, args.cpu_seconds) if fdk
                               else cpu_baseline(w, vol_np, geom_np, args.cpu_seconds))
    if args.workload == DEFAULT_WORKLOAD and not args.no_calibration:
        # the metric's second half: seconds to 0.1 px, learned vs plain start, per solver
        del vol, geom, upstream
        torch.cuda.empty_cache()
        out["calibration"] = calibration_time(args, torch, device)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Pin the CPU oracle: golden fixtures, invariants, closed forms, finite
differences and dot-product identities.  CPU only.

Tolerances follow the reference's own checks where they exist
(examples/check_forward_operator.py:167-191, examples/check_operator.py:155-173).
"""

import math

import numpy as np
import pytest

import cases
import oracle


def fp_golden():
    g = cases.load_golden()["fp_case"]
    return (np.array(g["volume"]), [np.array(x) for x in g["geometry"]], np.array(g["upstream"]), g)


def fp_call(vol, geom, g, mode="f64", traces=False):
    return oracle.fp_forward(vol, geom, g["pixel_size"], g["U"], g["V"], g["step_size"],
                             g["voxel_size"], tuple(g["offset"]), mode=mode, traces=traces)


# --------------------------------------------------------------------------- forward

def test_fp_oracle_matches_golden_scalar_restatement():
    vol, geom, _, g = fp_golden()
    out = fp_call(vol, geom, g)
    np.testing.assert_allclose(out, np.array(g["expected_projection_f64"]), rtol=1e-12, atol=1e-12)


def test_fp_oracle_invariants():
    """constant volume -> count*step, zero -> 0, linearity (check_forward_operator.py:164-173)."""
    vol, geom, _, g = fp_golden()
    out, (entry, count, _) = fp_call(vol, geom, g, traces=True)
    assert np.all(count > 0) and np.all(entry >= -1)
    const = fp_call(np.ones_like(vol), geom, g)
    np.testing.assert_allclose(const, count * g["step_size"], rtol=1e-12, atol=1e-12)
    assert np.all(fp_call(np.zeros_like(vol), geom, g) == 0)
    other = vol[:, :, ::-1].copy()
    comb = fp_call(1.3 * vol - 0.4 * other, geom, g)
    np.testing.assert_allclose(comb, 1.3 * out - 0.4 * fp_call(other, geom, g), rtol=1e-12, atol=1e-12)


def test_fp_oracle_f32_schedule_within_reference_tolerance():
    vol, geom, _, g = fp_golden()
    f64 = fp_call(vol, geom, g)
    f32 = fp_call(vol, geom, g, mode="f32")
    np.testing.assert_allclose(f32, f64, rtol=5e-5, atol=5e-5)


def test_fp_oracle_constant_volume_f32_counts_match_accumulated_schedule():
    """The FP32 schedule's sample count equals a float32 replay of t += step."""
    vol, geom, _, g = fp_golden()
    const32 = fp_call(np.ones_like(vol), geom, g, mode="f32")
    _, (_, count, _) = fp_call(vol, geom, g, traces=True)
    # counts agree with the ideal FP64 count on this fixture (no boundary ties)
    np.testing.assert_allclose(const32, count * np.float32(g["step_size"]), rtol=2e-6)


def _fd_geometry_check(mode_geometry, vol, geom, g, ray, weight, eps, rel, abs_):
    a, u, v = ray
    grad = np.zeros((len(geom[0]), g["U"], g["V"]))
    grad[a, u, v] = weight
    gvol, ggeo = oracle.fp_backward(grad, vol, geom, g["pixel_size"], g["step_size"], g["voxel_size"],
                                    tuple(g["offset"]), mode="f64")
    _, tr0 = fp_call(vol, geom, g, traces=True)
    base = (tr0[0][ray], tr0[1][ray], tr0[2][ray])
    checked = []
    for k in mode_geometry:
        plus = [x.copy() for x in geom]
        minus = [x.copy() for x in geom]
        plus[k][a] += eps
        minus[k][a] -= eps
        pp, trp = fp_call(vol, plus, g, traces=True)
        pm, trm = fp_call(vol, minus, g, traces=True)
        if (trp[0][ray], trp[1][ray], trp[2][ray]) != base or (trm[0][ray], trm[1][ray], trm[2][ray]) != base:
            continue
        fd = weight * (pp[ray] - pm[ray]) / (2 * eps)
        assert math.isclose(ggeo[k, a], fd, rel_tol=rel, abs_tol=abs_), (k, ggeo[k, a], fd)
        checked.append(k)
    return gvol, checked


def test_fp_oracle_geometry_gradients_finite_differences():
    """Central FD with trace guards (check_forward_operator.py:150-193); all eight
    geometry parameters are checked wherever the trace stays fixed."""
    vol, geom, _, g = fp_golden()
    gvol, checked = _fd_geometry_check(range(8), vol, geom, g, cases.FP_SELECTED_RAY,
                                       cases.FP_SELECTED_WEIGHT, 1e-5, 2e-6, 1e-7)
    # the reference's own selection (src_z, det_u0, det_v0) must be among the checked ones
    assert {7, 3, 4} <= set(checked), checked
    assert len(checked) >= 6
    # volume entry with the largest gradient
    idx = np.unravel_index(np.argmax(np.abs(gvol)), gvol.shape)
    plus, minus = vol.copy(), vol.copy()
    plus[idx] += 1e-5
    minus[idx] -= 1e-5
    fd = cases.FP_SELECTED_WEIGHT * (fp_call(plus, geom, g)[cases.FP_SELECTED_RAY]
                                     - fp_call(minus, geom, g)[cases.FP_SELECTED_RAY]) / 2e-5
    assert math.isclose(gvol[idx], fd, rel_tol=2e-6, abs_tol=1e-7)


@pytest.mark.parametrize("mode,rtol", [("f64", 1e-12), ("f32", 2e-5)])
def test_fp_oracle_adjoint_dot_product(mode, rtol):
    """<P x, y> == <x, P^T y>: the volume gradient is the exact transpose."""
    rng = np.random.default_rng(7)
    vol, geom, _, g = fp_golden()
    x = rng.random(vol.shape)
    y = rng.standard_normal((2, g["U"], g["V"]))
    px = fp_call(x, geom, g, mode=mode)
    gvol, _ = oracle.fp_backward(y, x, geom, g["pixel_size"], g["step_size"], g["voxel_size"],
                                 tuple(g["offset"]), mode=mode)
    lhs = float(np.sum(px.astype(np.float64) * y))
    rhs = float(np.sum(gvol.astype(np.float64) * x))
    assert math.isclose(lhs, rhs, rel_tol=rtol)


def test_fp_oracle_matches_gaussian_line_integrals():
    """Closed form: line integrals of Gaussian blobs along the source->pixel rays.
    Pins the ray geometry (rotation sense, source offsets, detector centres,
    volume offsets); a sign error in any of them moves a blob by several sigma."""
    D, H, W = 64, 68, 72
    vs = 1.0
    off = (1.5, -2.0, 0.75)
    A, U, V, pix = 6, 84, 80, 1.6
    rng = np.random.default_rng(3)
    geom = [np.linspace(0.3, 5.5, A), np.full(A, 160.0), np.full(A, 95.0),
            (U - 1) / 2 + rng.uniform(-2, 2, A), (V - 1) / 2 + rng.uniform(-2, 2, A),
            rng.uniform(-3, 3, A), rng.uniform(-3, 3, A), rng.uniform(-3, 3, A)]
    centres = np.array([[6.0, -4.0, 3.0], [-7.0, 5.0, -4.0], [2.0, 8.0, 6.0]])
    sig = [6.0, 5.5, 5.0]
    amp = [1.0, 0.7, 0.5]
    xs = (np.arange(W) - W / 2) * vs + off[0]
    ys = (np.arange(H) - H / 2) * vs + off[1]
    zs = (np.arange(D) - D / 2) * vs + off[2]
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    vol = np.zeros((D, H, W))
    for c, s, a_ in zip(centres, sig, amp):
        vol += a_ * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / (2 * s * s))
    out = oracle.fp_forward(vol, geom, pix, U, V, 0.25, vs, off, mode="f64")
    exact = oracle.gaussian_line_integrals(centres, sig, amp, geom, pix, U, V)
    err = cases.rel_l2(out, exact)
    assert err < 4e-3, err
    # the same check with the source offsets mirrored fails by a wide margin
    bad = list(geom)
    bad[5] = -geom[5]
    bad[7] = -geom[7]
    assert cases.rel_l2(oracle.fp_forward(vol, bad, pix, U, V, 0.25, vs, off, mode="f64"), exact) > 5 * err


def test_fp_oracle_sample_count_matches_constant_volume():
    """S from the count-only pass equals sum(constant projection)/step."""
    g = cases.C1
    geom = cases.c1_geometry()
    S, hits = oracle.fp_count_samples(g["shape"], geom, g["pixel"], g["U"], g["V"], g["step"], g["voxel"])
    const = oracle.fp_forward(np.ones(g["shape"], np.float32), geom, g["pixel"], g["U"], g["V"],
                              g["step"], g["voxel"], mode="f32")
    assert hits > 0
    assert S == int(round(float(np.sum(const.astype(np.float64))) / g["step"]))


# --------------------------------------------------------------------------- backprojector

def bp_golden():
    g = cases.load_golden()["bp_case"]
    return (np.array(g["projections"]), [np.array(x) for x in g["geometry"]], np.array(g["upstream"]), g)


def bp_call(proj, geom, g, mode="f64", chunk=0):
    return oracle.bp_forward(proj, geom, g["pixel_size"], tuple(g["volume_shape"]), g["voxel_size"],
                             tuple(g["offset"]), chunk=chunk, mode=mode)


def test_bp_oracle_matches_golden_scalar_restatement():
    proj, geom, up, g = bp_golden()
    out = bp_call(proj, geom, g)
    np.testing.assert_allclose(out, np.array(g["expected_volume_f64"]), rtol=1e-12, atol=1e-12)
    assert math.isclose(float(np.sum(out * up)), g["expected_weighted_loss_f64"], rel_tol=1e-12, abs_tol=1e-12)


def test_bp_oracle_finite_differences_all_nine_families():
    """Two entries per family with the largest |gradient|, eps 1e-6, rel 2e-6 / abs 1e-7
    (check_operator.py:150-177)."""
    proj, geom, up, g = bp_golden()
    gproj, ggeo = oracle.bp_backward(up, proj, geom, g["pixel_size"], g["voxel_size"], tuple(g["offset"]),
                                     mode="f64")

    def loss(p, gm):
        return math.fsum((bp_call(p, gm, g) * up).ravel().tolist())

    eps = 1e-6
    families = [gproj] + [ggeo[k] for k in range(8)]
    for fam, grad in enumerate(families):
        for flat in np.argsort(-np.abs(grad).ravel())[:2]:
            idx = np.unravel_index(flat, grad.shape)
            if fam == 0:
                pp, pm = proj.copy(), proj.copy()
                pp[idx] += eps
                pm[idx] -= eps
                fd = (loss(pp, geom) - loss(pm, geom)) / (2 * eps)
            else:
                gp = [x.copy() for x in geom]
                gm = [x.copy() for x in geom]
                gp[fam - 1][idx] += eps
                gm[fam - 1][idx] -= eps
                fd = (loss(proj, gp) - loss(proj, gm)) / (2 * eps)
            assert math.isclose(grad[idx], fd, rel_tol=2e-6, abs_tol=1e-7), (fam, idx, grad[idx], fd)


def test_bp_oracle_f32_within_reference_tolerance():
    proj, geom, up, g = bp_golden()
    np.testing.assert_allclose(bp_call(proj, geom, g, mode="f32"), bp_call(proj, geom, g), rtol=2e-5, atol=2e-5)
    g64 = oracle.bp_backward(up, proj, geom, g["pixel_size"], g["voxel_size"], tuple(g["offset"]), mode="f64")
    g32 = oracle.bp_backward(up, proj, geom, g["pixel_size"], g["voxel_size"], tuple(g["offset"]), mode="f32")
    for a32, a64 in zip(g32, g64):
        np.testing.assert_allclose(a32, a64, rtol=3e-4, atol=3e-5)


def test_bp_oracle_linear_projection_closed_form():
    """Projections linear in (u,v) are interpolated exactly, so each voxel receives
    sum_a w_a (alpha*u_a + beta*v_a + gamma) with (u_a, v_a, w_a) from intersecting the
    source->voxel line with the detector plane in world vectors."""
    A, U, V, pix = 5, 64, 56, 1.3
    shape, vs, off = (6, 7, 8), 1.1, (0.4, -0.3, 0.2)
    rng = np.random.default_rng(11)
    geom = [np.linspace(0.2, 4.9, A), rng.uniform(150, 170, A), rng.uniform(90, 100, A),
            (U - 1) / 2 + rng.uniform(-1, 1, A), (V - 1) / 2 + rng.uniform(-1, 1, A),
            rng.uniform(-2, 2, A), rng.uniform(-2, 2, A), rng.uniform(-2, 2, A)]
    al, be, ga = 0.013, -0.021, 0.7
    uu, vv = np.meshgrid(np.arange(U), np.arange(V), indexing="ij")
    proj = np.broadcast_to(al * uu + be * vv + ga, (A, U, V)).copy()
    out = oracle.bp_forward(proj, geom, pix, shape, vs, off, mode="f64")
    D, H, W = shape
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    pts = np.stack([(x - W / 2) * vs + off[0], (y - H / 2) * vs + off[1], (z - D / 2) * vs + off[2]], -1)
    ui, vi, w = oracle.point_detector_coordinates(pts.reshape(-1, 3), geom, pix)
    assert np.all((ui >= 0) & (ui < U - 1) & (vi >= 0) & (vi < V - 1))
    expect = np.sum(w * (al * ui + be * vi + ga), axis=0).reshape(shape)
    np.testing.assert_allclose(out, expect, rtol=1e-11, atol=1e-11)


def test_bp_oracle_chunking_is_a_regrouping():
    proj, geom, _, g = bp_golden()
    full = bp_call(proj, geom, g)
    for chunk in (1, 2):
        np.testing.assert_allclose(bp_call(proj, geom, g, chunk=chunk), full, rtol=1e-13, atol=1e-13)


# --------------------------------------------------------------------------- FDK

def test_fdk_filter_fft_equals_direct_convolution():
    rng = np.random.default_rng(5)
    for U, V, A in ((48, 40, 3), (37, 5, 2)):
        p = rng.random((A, U, V))
        sod, sid = rng.uniform(90, 110, A), rng.uniform(150, 170, A)
        a = oracle.fdk_filter(p, 0.8, sod, sid, 0.8)
        b = oracle.fdk_filter_direct(p, 0.8, sod, sid, 0.8)
        np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-11 * np.abs(a).max())


def test_fdk_ramp_response_matches_closed_form_ramp():
    """The discrete Ram-Lak kernel h[0]=1/(4p^2), h[odd n]=-1/(pi n p)^2 has the DTFT
    |f|/p^2 (Kak & Slaney, eq. 3.61); truncated to the padded length the response
    differs by O(1/pad)."""
    for U, pix in ((48, 1.0), (512, 0.8)):
        pad = oracle.fdk_pad(U)
        f = np.fft.fftfreq(pad)
        ramp_only = oracle.ramp_hann_response(U, pix, alpha=1.0) / np.where(
            np.abs(f) <= 0.5, 0.5 + 0.5 * np.cos(2 * np.pi * f), 1.0).clip(1e-300)
        inner = np.abs(f) < 0.45
        err = np.abs(ramp_only[inner] - np.abs(f[inner]) / pix ** 2).max() * pix ** 2
        assert err < 4.0 / pad, (U, err)


def test_fdk_delta_response():
    A, U, V, pix = 1, 40, 3, 0.9
    sod, sid = 100.0, 170.0
    p = np.zeros((A, U, V))
    k = 17
    p[0, k, 1] = 1.0
    out = oracle.fdk_filter(p, pix, sod, sid, 0.8)
    taps = oracle.filter_taps(U, pix, 0.8)
    u = (k - (U - 1) / 2) * pix
    v = (1 - (V - 1) / 2) * pix
    pre = sid / math.sqrt(sid ** 2 + u * u + v * v)
    expect = taps[np.arange(U) - k + U - 1] * pre * pix * sid / sod
    np.testing.assert_allclose(out[0, :, 1], expect, rtol=1e-12, atol=1e-15)
    assert np.all(out[0, :, 0] == 0) and np.all(out[0, :, 2] == 0)


def sphere_projections(A, U, V, pix, sod, sid, radius):
    """Analytic line integrals of a centred unit-density sphere (examples/fdk.py:37-51)."""
    u = (np.arange(U) - (U - 1) / 2) * pix
    v = (np.arange(V) - (V - 1) / 2) * pix
    r2 = u[:, None] ** 2 + v[None, :] ** 2
    d2 = sod ** 2 * r2 / (sid ** 2 + r2)
    view = 2.0 * np.sqrt(np.clip(radius ** 2 - d2, 0, None))
    return np.broadcast_to(view, (A, U, V)).copy()


def test_fdk_reconstructs_centred_sphere():
    """Known answer: FDK of a centred unit sphere is ~1 inside and ~0 well outside."""
    A, U, V, pix, sod, sid = 90, 96, 96, 1.0, 500.0, 1000.0
    radius = 10.0
    shape, vs = (48, 48, 48), 0.5
    proj = sphere_projections(A, U, V, pix, sod, sid, radius)
    ang = np.arange(A) * 2 * math.pi / A
    geom = [ang, np.full(A, sid), np.full(A, sod), np.full(A, (U - 1) / 2), np.full(A, (V - 1) / 2),
            np.zeros(A), np.zeros(A), np.zeros(A)]
    vol = oracle.reconstruct_fdk(proj, geom, pix, shape, vs)
    D, H, W = shape
    z, y, x = np.meshgrid(np.arange(D) - D / 2, np.arange(H) - H / 2, np.arange(W) - W / 2, indexing="ij")
    r = np.sqrt(x * x + y * y + z * z) * vs
    inside = vol[r < 0.7 * radius]
    outside = vol[r > 1.3 * radius]
    assert abs(inside.mean() - 1.0) < 0.03, inside.mean()
    assert outside.mean() < 0.03


def test_fdk_chunk_invariance():
    rng = np.random.default_rng(2)
    A, U, V = 10, 20, 12
    proj = rng.random((A, U, V))
    geom = [np.arange(A) * 0.6, np.full(A, 150.0), np.full(A, 90.0), np.full(A, (U - 1) / 2),
            np.full(A, (V - 1) / 2), np.zeros(A), np.zeros(A), np.zeros(A)]
    ref = oracle.reconstruct_fdk(proj, geom, 1.0, (6, 8, 8), 1.0, angle_chunk_size=A)
    for c in (1, 3, 7):
        np.testing.assert_allclose(oracle.reconstruct_fdk(proj, geom, 1.0, (6, 8, 8), 1.0, angle_chunk_size=c),
                                   ref, rtol=1e-12, atol=1e-12)


def test_preprocess_oracle_recovers_line_integrals_and_transposes():
    """Known answer: frames made from line integrals p by Beer-Lambert with dark and flat
    fields give back p (transposed from [A,V,U] to [A,U,V]); fully absorbed or dead pixels
    clamp at -log(t_min)."""
    rng = np.random.default_rng(3)
    A, V, U = 2, 5, 7
    p = rng.uniform(0.0, 4.0, (A, U, V))
    dark = rng.uniform(90, 110, (V, U))
    flat = rng.uniform(3000, 4000, (V, U))
    raw = dark + (flat - dark) * np.exp(-p.transpose(0, 2, 1))
    got = oracle.preprocess(raw, dark, flat, "AVU")
    assert got.shape == (A, U, V)
    assert np.allclose(got, p, rtol=1e-12, atol=1e-12)
    raw2 = raw.copy()
    raw2[0, 0, 0] = dark[0, 0] - 5            # below dark: transmission clamps
    flat2 = flat.copy()
    flat2[1, 2] = dark[1, 2]                    # dead pixel
    got2 = oracle.preprocess(raw2, dark, flat2, "AVU", min_transmission=1e-4)
    assert got2[0, 0, 0] == -np.log(1e-4) and got2[0, 2, 1] == -np.log(1e-4)
    assert np.array_equal(oracle.preprocess(raw, layout="AUV", log=False), raw)


def test_parker_weights_sum_to_one_over_conjugate_rays():
    """Every line measured twice in a short scan gets total weight one: the conjugate of the ray
    (beta, gamma) is (beta + pi + 2 gamma, -gamma) in this geometry (checked from the ray
    definition itself), and the two Parker weights add up to 1."""
    sod, sid, pix, U = 500.0, 1000.0, 1.0, 201
    # the conjugate relation, from the source/detector construction of the projector
    beta, up = 0.3, 60.0
    s, c = math.sin(beta), math.cos(beta)
    O = np.array([-sod * s, sod * c])
    T = np.array([up * c - (sod - sid) * s, up * s + (sod - sid) * c])
    d = (T - O) / np.linalg.norm(T - O)
    P2 = O - 2 * O.dot(d) * d
    beta2 = math.atan2(-P2[0], P2[1])
    gamma = math.atan(up / sid)
    assert abs(((beta2 - beta) - (math.pi + 2 * gamma)) % (2 * math.pi)) < 1e-9
    # redundancy of the weights: evaluate each ray and its conjugate exactly (a three-view scan
    # [0, beta, range] keeps delta and places the view of interest at beta)
    delta = math.atan(((U - 1) / 2) * pix / sid)
    rng_ = math.pi + 2 * delta
    g = np.arctan((np.arange(U) - (U - 1) / 2) * pix / sid)
    weight = lambda b, col: oracle.parker_weights(np.array([0.0, b, rng_]), U, pix, sid)[1, col]
    w_all = oracle.parker_weights(np.linspace(0.0, rng_, 361), U, pix, sid)
    assert w_all.min() >= 0.0 and w_all.max() <= 1.0 and np.all(w_all[0] == 0.0)
    rs = np.random.default_rng(0)
    checked = 0
    for _ in range(400):
        b, u = rs.uniform(0.0, rng_), int(rs.integers(U))
        b2 = b + math.pi + 2 * g[u]
        if b2 > rng_:
            b2 -= 2 * math.pi
        if not (0.0 <= b2 <= rng_):
            assert weight(b, u) == 1.0                # measured once: full weight
            continue
        u2 = (U - 1) - u                             # the column of fan angle -gamma
        assert abs(weight(b, u) + weight(b2, u2) - 1.0) < 1e-9
        checked += 1
    assert checked > 30


def test_short_scan_fdk_reconstructs_centred_sphere():
    A, U, V, pix, sod, sid = 120, 96, 96, 1.0, 500.0, 1000.0
    radius = 10.0
    delta = math.atan(((U - 1) / 2) * pix / sid)
    ang = np.linspace(0.0, math.pi + 2 * delta, A)
    proj = sphere_projections(A, U, V, pix, sod, sid, radius)
    geom = [ang, np.full(A, sid), np.full(A, sod), np.full(A, (U - 1) / 2), np.full(A, (V - 1) / 2),
            np.zeros(A), np.zeros(A), np.zeros(A)]
    vol = oracle.reconstruct_fdk(proj, geom, pix, (48, 48, 48), 0.5, short_scan=True)
    z, y, x = np.meshgrid(*(np.arange(48) - 24,) * 3, indexing="ij")
    r = np.sqrt(x * x + y * y + z * z) * 0.5
    assert abs(vol[r < 7.0].mean() - 1.0) < 0.03
    assert vol[r > 13.0].mean() < 0.03


def test_offset_preweight_at_the_centre_equals_the_reference_preweight():
    rng = np.random.default_rng(9)
    A, U, V = 3, 20, 14
    p = rng.random((A, U, V))
    ref = oracle.fdk_filter(p, 0.7, 300.0, 500.0)
    got = oracle.fdk_filter(p, 0.7, 300.0, 500.0, u0=np.full(A, (U - 1) / 2), v0=np.full(A, (V - 1) / 2))
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)

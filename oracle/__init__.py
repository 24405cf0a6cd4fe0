"""CPU oracle for the cone-beam CT operators -- TEST INFRASTRUCTURE ONLY.

This package is the checker: ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` use it to check (or time beside) the HIP
library.  The product (``cbct_backprojector``) never imports it.

Contents and the reference code each part restates (paths relative to the
reference repository):

* ``fp_forward`` / ``fp_backward`` / ``fp_count_samples`` -- C library
  ``liboracle.so`` (``cbct_oracle.c``, ``oracle_core.inc``) restating
  ``csrc/src_cbct_forwardprojector_cuda.cu:55-528``;
* ``bp_forward`` / ``bp_backward`` -- same library, restating
  ``csrc/src_cbct_backprojector_cuda.cu:15-295``;
* ``fdk_filter`` -- numpy FP64 restatement of ``cbct_backprojector/filtering.py:54-104``
  through the FFT, and ``fdk_filter_direct``, the same filter as a direct
  linear convolution (a second algorithm used to pin the first);
* ``reconstruct_fdk`` -- ``cbct_backprojector/filtering.py:107-198``;
* ``ray_integral_py`` / ``bp_scalar_py`` -- small pure-Python FP64 restatements
  (``examples/check_forward_operator.py:61-105``,
  ``examples/check_operator.py:97-138``) used on tiny cases only;
* ``ellipsoid_chords`` -- analytic line integrals through ellipsoids (a
  closed form that pins the ray geometry independently of the restatement).

Every function of the oracle is pinned by tests in ``tests/test_oracle.py``
(invariants, closed forms, finite differences, dot-product identities and
the reference's own fixtures); none is "parity unpinned".

Two arithmetic modes: ``"f32"`` is the FP32 reference schedule (float
arithmetic with the operation order fixed in ``oracle_core.inc`` and the step
accumulated in float as ``csrc/src_cbct_forwardprojector_cuda.cu:321``), and
``"f64"`` is the FP64 oracle (double arithmetic, sample k at ``t_min + k*step``,
as ``examples/check_forward_operator.py:97-99``).
"""

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

GEOMETRY_NAMES = ("angles", "sid", "sod", "det_u0", "det_v0", "src_x", "src_y", "src_z")


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
        _declare(_LIB)
    return _LIB


def _declare(lib):
    P = ctypes.c_void_p
    I = ctypes.c_int
    for sfx, R in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
        f = getattr(lib, f"orc_fp_forward_{sfx}")
        f.argtypes = [P, I, I, I] + [P] * 8 + [I, R, I, I, R, R, R, R, R, P, P, P, P, I]
        f.restype = None
        f = getattr(lib, f"orc_fp_count_samples_{sfx}")
        f.argtypes = [I, I, I] + [P] * 8 + [I, R, I, I, R, R, R, R, R, P, I]
        f.restype = ctypes.c_int64
        f = getattr(lib, f"orc_fp_backward_{sfx}")
        f.argtypes = [P, P, I, I, I] + [P] * 8 + [I, R, I, I, R, R, R, R, R, P, P, I]
        f.restype = None
        f = getattr(lib, f"orc_bp_forward_{sfx}")
        f.argtypes = [P, I, I, I] + [P] * 8 + [R, I, I, I, R, R, R, R, I, P, I]
        f.restype = None
        f = getattr(lib, f"orc_bp_backward_{sfx}")
        f.argtypes = [P, P, I, I, I] + [P] * 8 + [R, I, I, I, R, R, R, R, P, P, I]
        f.restype = None
    lib.orc_max_threads.restype = I


def max_threads():
    return int(_lib().orc_max_threads())


def _dt(mode):
    if mode not in ("f32", "f64"):
        raise ValueError("mode must be 'f32' or 'f64'")
    return np.float32 if mode == "f32" else np.float64


def _arr(x, dt):
    return np.ascontiguousarray(np.asarray(x, dtype=dt))


def _geom(geometry, dt):
    if isinstance(geometry, dict):
        geometry = [geometry[name] for name in GEOMETRY_NAMES]
    arrays = [_arr(g, dt).reshape(-1) for g in geometry]
    if len(arrays) != 8 or len({a.size for a in arrays}) != 1:
        raise ValueError("geometry needs 8 arrays of equal length")
    return arrays


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def check_step_resolution(volume_shape, step_size, voxel_size):
    """The library's host check: a step below 2^-22 (volume diagonal + 1) is refused (every view
    would be degenerate, since no march could follow it in FP32)."""
    D, H, W = volume_shape
    diag = float(voxel_size) * math.sqrt(float(W) ** 2 + float(H) ** 2 + float(D) ** 2)
    R = float(np.float32(0.5 * diag))
    if float(step_size) < (2.0 * R + 1.0) * 2.0 ** -22:
        raise ValueError("step_size is below the FP32 resolution of the ray parameter over the volume")


def fp_forward(volume, geometry, pixel_size, U, V, step_size, voxel_size,
               offset=(0.0, 0.0, 0.0), mode="f32", traces=False, threads=None):
    """Ray integrals [A,U,V]; with ``traces`` also (entry_axis, count, cell_hash) per ray."""
    dt = _dt(mode)
    vol = _arr(volume, dt)
    g = _geom(geometry, dt)
    A = g[0].size
    D, H, W = vol.shape
    check_step_resolution((D, H, W), step_size, voxel_size)
    out = np.zeros((A, U, V), dt)
    tr = (np.zeros((A, U, V), np.int32), np.zeros((A, U, V), np.int64), np.zeros((A, U, V), np.uint64)) if traces else None
    fn = getattr(_lib(), f"orc_fp_forward_{mode}")
    fn(_ptr(vol), D, H, W, *[_ptr(x) for x in g], A, pixel_size, U, V, step_size, voxel_size,
       offset[0], offset[1], offset[2], _ptr(out),
       _ptr(tr[0]) if tr else None, _ptr(tr[1]) if tr else None, _ptr(tr[2]) if tr else None,
       threads or max_threads())
    return (out, tr) if traces else out


def fp_count_samples(volume_shape, geometry, pixel_size, U, V, step_size, voxel_size,
                     offset=(0.0, 0.0, 0.0), mode="f32", threads=None):
    """(S, hitting rays): the exact number of trilinear samples of the forward schedule."""
    dt = _dt(mode)
    g = _geom(geometry, dt)
    D, H, W = volume_shape
    check_step_resolution((D, H, W), step_size, voxel_size)
    hits = ctypes.c_int64(0)
    fn = getattr(_lib(), f"orc_fp_count_samples_{mode}")
    S = fn(D, H, W, *[_ptr(x) for x in g], g[0].size, pixel_size, U, V, step_size, voxel_size,
           offset[0], offset[1], offset[2], ctypes.cast(ctypes.pointer(hits), ctypes.c_void_p),
           threads or max_threads())
    return int(S), int(hits.value)


def fp_backward(grad, volume, geometry, pixel_size, step_size, voxel_size,
                offset=(0.0, 0.0, 0.0), need_volume=True, mode="f32", threads=None):
    """(gvol or None, [8, A] geometry gradients) of sum(grad * fp_forward)."""
    dt = _dt(mode)
    gr = _arr(grad, dt)
    vol = _arr(volume, dt)
    g = _geom(geometry, dt)
    A, U, V = gr.shape
    D, H, W = vol.shape
    check_step_resolution((D, H, W), step_size, voxel_size)
    gvol = np.zeros_like(vol) if need_volume else None
    ggeo = np.zeros((8, A), dt)
    fn = getattr(_lib(), f"orc_fp_backward_{mode}")
    fn(_ptr(gr), _ptr(vol), D, H, W, *[_ptr(x) for x in g], A, pixel_size, U, V, step_size,
       voxel_size, offset[0], offset[1], offset[2], _ptr(gvol) if need_volume else None,
       _ptr(ggeo), threads or max_threads())
    return gvol, ggeo


def bp_forward(projections, geometry, pixel_size, volume_shape, voxel_size,
               offset=(0.0, 0.0, 0.0), chunk=0, mode="f32", threads=None):
    """Weighted backprojection [D,H,W]; ``chunk`` groups views like reconstruct_fdk."""
    dt = _dt(mode)
    pr = _arr(projections, dt)
    g = _geom(geometry, dt)
    A, U, V = pr.shape
    D, H, W = volume_shape
    out = np.zeros((D, H, W), dt)
    fn = getattr(_lib(), f"orc_bp_forward_{mode}")
    fn(_ptr(pr), A, U, V, *[_ptr(x) for x in g], pixel_size, D, H, W, voxel_size,
       offset[0], offset[1], offset[2], int(chunk), _ptr(out), threads or max_threads())
    return out


def bp_backward(grad_volume, projections, geometry, pixel_size, voxel_size,
                offset=(0.0, 0.0, 0.0), mode="f32", threads=None):
    """(gproj [A,U,V], [8, A] geometry gradients) of sum(grad_volume * bp_forward)."""
    dt = _dt(mode)
    gv = _arr(grad_volume, dt)
    pr = _arr(projections, dt)
    g = _geom(geometry, dt)
    A, U, V = pr.shape
    D, H, W = gv.shape
    gproj = np.zeros_like(pr)
    ggeo = np.zeros((8, A), dt)
    fn = getattr(_lib(), f"orc_bp_backward_{mode}")
    fn(_ptr(gv), _ptr(pr), A, U, V, *[_ptr(x) for x in g], pixel_size, D, H, W, voxel_size,
       offset[0], offset[1], offset[2], _ptr(gproj), _ptr(ggeo), threads or max_threads())
    return gproj, ggeo


# ---------------------------------------------------------------------------
# FDK filter and reconstruction (cbct_backprojector/filtering.py:54-198)
# ---------------------------------------------------------------------------

def _per_view(value, A):
    arr = np.asarray(value, np.float64).reshape(-1)
    if arr.size == 1:
        arr = np.full(A, float(arr[0]))
    if arr.size != A:
        raise ValueError("distance must be a scalar or have one value per view")
    return arr


def fdk_pad(U):
    """Zero-padded length along U: 1 << bit_length(2U-1) (filtering.py:87)."""
    return 1 << (2 * U - 1).bit_length()


def ramp_hann_response(U, pixel_size, alpha=0.8):
    """Frequency response Re(FFT(ramp)) * Hann over the padded length (filtering.py:88-101)."""
    pad = fdk_pad(U)
    idx = np.arange(pad)
    n = np.where(idx > pad // 2, idx - pad, idx)
    h = np.zeros(pad)
    h[0] = 1.0 / (4.0 * pixel_size ** 2)
    odd = n % 2 != 0
    h[odd] = -1.0 / (math.pi * n[odd] * pixel_size) ** 2
    ramp = np.fft.fft(h).real
    freq = np.fft.fftfreq(pad)
    hann = np.where(np.abs(freq) <= alpha * 0.5, 0.5 + 0.5 * np.cos(2.0 * math.pi * freq / alpha), 0.0)
    return ramp * hann


def filter_taps(U, pixel_size, alpha=0.8):
    """h_eff[d] for d = -(U-1) .. U-1: the linear-convolution kernel the padded FFT applies."""
    pad = fdk_pad(U)
    hc = np.fft.ifft(ramp_hann_response(U, pixel_size, alpha)).real
    d = np.arange(-(U - 1), U)
    return hc[d % pad]


def parker_weights(angles, U, pixel_size, sid, u0=None):
    """Parker short-scan weights [A, U] (Parker 1982; Kak & Slaney ch. 3) for fan angle
    gamma = atan(u_p / sid) of each column, beta = angle - angle_0 and delta = (range - pi)/2;
    the conjugate of the ray (beta, gamma) in this geometry is (beta + pi + 2 gamma, -gamma)."""
    ang = np.asarray(angles, np.float64).reshape(-1)
    A = ang.size
    beta = (ang - ang[0])[:, None]
    delta = max((ang[-1] - ang[0] - math.pi) / 2.0, 1e-9)
    cu = np.full(A, (U - 1) / 2.0) if u0 is None else np.asarray(u0, np.float64).reshape(-1)
    g = np.arctan((np.arange(U)[None, :] - cu[:, None]) * pixel_size / _per_view(sid, A)[:, None])
    w = np.zeros((A, U))
    r1 = (beta >= 0) & (beta < 2 * delta - 2 * g)
    r2 = (beta >= 2 * delta - 2 * g) & (beta <= math.pi - 2 * g)
    r3 = (beta > math.pi - 2 * g) & (beta <= math.pi + 2 * delta)
    with np.errstate(divide="ignore", invalid="ignore"):
        w = np.where(r1, np.sin(math.pi / 4 * beta / (delta - g)) ** 2, w)
        w = np.where(r2, 1.0, w)
        w = np.where(r3, np.sin(math.pi / 4 * (math.pi + 2 * delta - beta) / (delta + g)) ** 2, w)
    return w


def _preweight(projections, pixel_size, sod, sid, u0=None, v0=None, parker_angles=None):
    A, U, V = projections.shape
    sid = _per_view(sid, A)[:, None, None]
    cu = np.full(A, (U - 1) / 2.0) if u0 is None else np.asarray(u0, np.float64).reshape(-1)
    cv = np.full(A, (V - 1) / 2.0) if v0 is None else np.asarray(v0, np.float64).reshape(-1)
    u = (np.arange(U)[None, :] - cu[:, None]) * pixel_size
    v = (np.arange(V)[None, :] - cv[:, None]) * pixel_size
    dist = np.sqrt(sid ** 2 + u[:, :, None] ** 2 + v[:, None, :] ** 2)
    w = np.asarray(projections, np.float64) * (sid / dist)
    if parker_angles is not None:
        w = w * parker_weights(parker_angles, U, pixel_size, sid[:, 0, 0], u0)[:, :, None]
    return w


def fdk_filter(projections, pixel_size, sod, sid, alpha=0.8, u0=None, v0=None, parker_angles=None):
    """FP64 FFT restatement of fdk_filter (filtering.py:54-104); u0/v0 centre the preweight
    per view and parker_angles apply short-scan weights (extensions beyond the reference)."""
    p = np.asarray(projections, np.float64)
    A, U, V = p.shape
    w = _preweight(p, pixel_size, sod, sid, u0, v0, parker_angles)
    pad = fdk_pad(U)
    F = np.fft.fft(w, n=pad, axis=1) * ramp_hann_response(U, pixel_size, alpha)[None, :, None]
    out = np.fft.ifft(F, axis=1).real[:, :U, :]
    scale = pixel_size * (_per_view(sid, A) / _per_view(sod, A))
    return out * scale[:, None, None]


def fdk_filter_direct(projections, pixel_size, sod, sid, alpha=0.8):
    """The same filter as a direct linear convolution along U with filter_taps."""
    p = np.asarray(projections, np.float64)
    A, U, V = p.shape
    w = _preweight(p, pixel_size, sod, sid)
    taps = filter_taps(U, pixel_size, alpha)
    # out[u] = sum_u' h[u - u'] w[u']; taps index (u - u') + U - 1
    T = np.empty((U, U))
    for u in range(U):
        T[u] = taps[u - np.arange(U) + U - 1]
    out = np.einsum("ij,ajv->aiv", T, w)
    scale = pixel_size * (_per_view(sid, A) / _per_view(sod, A))
    return out * scale[:, None, None]


def reconstruct_fdk(projections, geometry, pixel_size, volume_shape, voxel_size,
                    offset=(0.0, 0.0, 0.0), alpha=0.8, angle_chunk_size=32, mode="f64",
                    short_scan=False, offset_preweight=False):
    """filter -> chunked weighted backprojection -> x pi/A -> clamp at 0 (filtering.py:182-198);
    short_scan: Parker weights and x (range / (A - 1)) instead of pi/A."""
    g = _geom(geometry, np.float64)
    A = g[0].size
    filt = fdk_filter(projections, pixel_size, g[2], g[1], alpha,
                      u0=g[3] if offset_preweight else None, v0=g[4] if offset_preweight else None,
                      parker_angles=g[0] if short_scan else None)
    vol = bp_forward(filt, g, pixel_size, volume_shape, voxel_size, offset,
                     chunk=angle_chunk_size, mode=mode)
    scale = (g[0][-1] - g[0][0]) / (A - 1) if short_scan else math.pi / A
    vol = vol * scale
    return np.maximum(vol, 0.0)


# ---------------------------------------------------------------------------
# Small pure-Python FP64 restatements (tiny cases only)
# ---------------------------------------------------------------------------

def ray_integral_py(volume, geometry, pixel_size, U, V, step_size, voxel_size, offset):
    """Scalar-loop ray integrals in FP64 (examples/check_forward_operator.py:61-105)."""
    vol = np.asarray(volume, np.float64)
    D, H, W = vol.shape
    ang, sid, sod, u0, v0, sx, sy, sz = [np.asarray(x, np.float64).reshape(-1) for x in
                                         (geometry if not isinstance(geometry, dict)
                                          else [geometry[k] for k in GEOMETRY_NAMES])]
    sizes = (W, H, D)
    lo = [(-n * 0.5 - 0.5) * voxel_size + o for n, o in zip(sizes, offset)]
    hi = [(n * 0.5 - 0.5) * voxel_size + o for n, o in zip(sizes, offset)]
    out = np.zeros((ang.size, U, V))
    for a in range(ang.size):
        c, s = math.cos(ang[a]), math.sin(ang[a])
        O = (-sod[a] * s + sx[a], sod[a] * c + sy[a], sz[a])
        n = sod[a] - sid[a]
        for u in range(U):
            up = (u - u0[a]) * pixel_size
            for v in range(V):
                vp = (v - v0[a]) * pixel_size
                T = (up * c - n * s, up * s + n * c, vp)
                r = [T[k] - O[k] for k in range(3)]
                L = math.sqrt(sum(x * x for x in r))
                if L <= 1e-8:
                    continue
                d = [x / L for x in r]
                tmin, tmax, ok = 0.0, math.inf, True
                for k in range(3):
                    if abs(d[k]) < 1e-12:
                        if not lo[k] <= O[k] <= hi[k]:
                            ok = False
                            break
                        continue
                    t1, t2 = (lo[k] - O[k]) / d[k], (hi[k] - O[k]) / d[k]
                    t1, t2 = min(t1, t2), max(t1, t2)
                    tmin, tmax = max(tmin, t1), min(tmax, t2)
                    if tmax < tmin:
                        ok = False
                        break
                if not ok:
                    continue
                count = math.floor((tmax - tmin) / step_size) + 1
                total = 0.0
                for k in range(count):
                    t = tmin + k * step_size
                    p = [(O[i] + t * d[i] - offset[i]) / voxel_size + sizes[i] * 0.5 for i in range(3)]
                    raw = [math.floor(x) for x in p]
                    fr = [p[i] - raw[i] for i in range(3)]
                    val = 0.0
                    for dz in (0, 1):
                        z = min(max(raw[2] + dz, 0), D - 1)
                        wz = fr[2] if dz else 1 - fr[2]
                        for dy in (0, 1):
                            y = min(max(raw[1] + dy, 0), H - 1)
                            wy = fr[1] if dy else 1 - fr[1]
                            for dx in (0, 1):
                                x = min(max(raw[0] + dx, 0), W - 1)
                                wx = fr[0] if dx else 1 - fr[0]
                                val += wx * wy * wz * vol[z, y, x]
                    total += val
                out[a, u, v] = total * step_size
    return out


def bp_scalar_py(projections, geometry, pixel_size, volume_shape, voxel_size, offset, weights=None):
    """Scalar FP64 weighted backprojection with math.fsum (examples/check_operator.py:97-138).

    Returns the volume, or sum(weights * volume) when ``weights`` is given."""
    proj = np.asarray(projections, np.float64)
    A, U, V = proj.shape
    ang, sid, sod, u0, v0, sx, sy, sz = [np.asarray(x, np.float64).reshape(-1) for x in
                                         (geometry if not isinstance(geometry, dict)
                                          else [geometry[k] for k in GEOMETRY_NAMES])]
    D, H, W = volume_shape
    vol = np.zeros((D, H, W))
    for z in range(D):
        vz = (z - D * 0.5) * voxel_size + offset[2]
        for y in range(H):
            vy = (y - H * 0.5) * voxel_size + offset[1]
            for x in range(W):
                vx = (x - W * 0.5) * voxel_size + offset[0]
                terms = []
                for a in range(A):
                    c, s = math.cos(ang[a]), math.sin(ang[a])
                    px, py = vx * c + vy * s, -vx * s + vy * c
                    dsu = sx[a] * c + sy[a] * s
                    dsn = -sx[a] * s + sy[a] * c
                    q = sod[a] + dsn - py
                    l = sid[a] + dsn
                    if q <= 1e-6 or l <= 1e-6:
                        continue
                    t = l / q
                    ui = (dsu + t * (px - dsu)) / pixel_size + u0[a]
                    vi = (sz[a] + t * (vz - sz[a])) / pixel_size + v0[a]
                    if not (0 <= ui < U - 1 and 0 <= vi < V - 1):
                        continue
                    iu, iv = math.floor(ui), math.floor(vi)
                    fu, fv = ui - iu, vi - iv
                    img = proj[a]
                    interp = ((1 - fu) * (1 - fv) * img[iu, iv] + fu * (1 - fv) * img[iu + 1, iv]
                              + (1 - fu) * fv * img[iu, iv + 1] + fu * fv * img[iu + 1, iv + 1])
                    terms.append(((sod[a] + dsn) / q) ** 2 * interp)
                vol[z, y, x] = math.fsum(terms)
    if weights is None:
        return vol
    return math.fsum((np.asarray(weights, np.float64) * vol).reshape(-1).tolist())


# ---------------------------------------------------------------------------
# Closed forms
# ---------------------------------------------------------------------------

def ellipsoid_chords(ellipsoids, geometry, pixel_size, U, V):
    """Exact line integrals of a sum of uniform ellipsoids along the source->pixel rays.

    ``ellipsoids`` is an [n, 11] array of (cx, cy, cz, ax, ay, az, qw, qx, qy, qz, density)
    in world coordinates (rotation given as a unit quaternion).  The ray of pixel (u,v)
    runs from the source O through the detector point T (the geometry convention of
    csrc/src_cbct_forwardprojector_cuda.cu:12-47), computed here in plain FP64 vectors.
    The chord is integrated over the whole line, so a detector placed beyond the object
    gives the integral over the object."""
    E = np.asarray(ellipsoids, np.float64).reshape(-1, 11)
    ang, sid, sod, u0, v0, sx, sy, sz = [np.asarray(x, np.float64).reshape(-1) for x in
                                         (geometry if not isinstance(geometry, dict)
                                          else [geometry[k] for k in GEOMETRY_NAMES])]
    A = ang.size
    out = np.zeros((A, U, V))
    uu = np.arange(U)[:, None] * np.ones((1, V))
    vv = np.ones((U, 1)) * np.arange(V)[None, :]
    for a in range(A):
        c, s = math.cos(ang[a]), math.sin(ang[a])
        ex = np.array([c, s, 0.0])          # detector u axis in world coordinates
        en = np.array([-s, c, 0.0])         # direction from isocentre towards the source
        ez = np.array([0.0, 0.0, 1.0])
        O = sod[a] * en + np.array([sx[a], sy[a], sz[a]])
        T = ((uu - u0[a]) * pixel_size)[..., None] * ex + (sod[a] - sid[a]) * en \
            + ((vv - v0[a]) * pixel_size)[..., None] * ez
        d = T - O
        d /= np.linalg.norm(d, axis=-1, keepdims=True)
        for e in E:
            R = _quat_to_matrix(e[6:10])
            axes = e[3:6]
            o_l = (R.T @ (O - e[0:3])) / axes
            d_l = (d @ R) / axes
            qa = np.sum(d_l * d_l, axis=-1)
            qb = 2.0 * np.sum(d_l * o_l, axis=-1)
            qc = np.dot(o_l, o_l) - 1.0
            disc = qb * qb - 4 * qa * qc
            chord = np.where(disc > 0, np.sqrt(np.maximum(disc, 0.0)) / qa, 0.0)
            out[a] += e[10] * chord
    return out


def _quat_to_matrix(q):
    w, x, y, z = np.asarray(q, np.float64) / np.linalg.norm(q)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def _pixel_rays(geometry, pixel_size, U, V, a):
    """Source position and unit directions [U,V,3] of view a in plain FP64 vector form:
    source at sod*n_hat + (sx,sy,sz), detector point at u_p*e_u + (sod-sid)*n_hat + v_p*e_z
    with e_u = (cos b, sin b, 0), n_hat = (-sin b, cos b, 0) (the convention stated in
    csrc/src_cbct_forwardprojector_cuda.cu:12-47)."""
    ang, sid, sod, u0, v0, sx, sy, sz = [np.asarray(x, np.float64).reshape(-1) for x in
                                         (geometry if not isinstance(geometry, dict)
                                          else [geometry[k] for k in GEOMETRY_NAMES])]
    c, s = math.cos(ang[a]), math.sin(ang[a])
    eu, en, ez = np.array([c, s, 0.0]), np.array([-s, c, 0.0]), np.array([0.0, 0.0, 1.0])
    O = sod[a] * en + np.array([sx[a], sy[a], sz[a]])
    up = (np.arange(U) - u0[a]) * pixel_size
    vp = (np.arange(V) - v0[a]) * pixel_size
    T = up[:, None, None] * eu + (sod[a] - sid[a]) * en + vp[None, :, None] * ez
    d = T - O
    return O, d / np.linalg.norm(d, axis=-1, keepdims=True)


def gaussian_line_integrals(centres, sigmas, amplitudes, geometry, pixel_size, U, V):
    """Exact integrals over the source->pixel lines of sum_j a_j exp(-|x-c_j|^2 / (2 s_j^2)):
    a_j * s_j * sqrt(2 pi) * exp(-d_j^2 / (2 s_j^2)), d_j the distance from c_j to the line."""
    A = np.asarray(geometry[0] if not isinstance(geometry, dict) else geometry["angles"]).size
    out = np.zeros((A, U, V))
    for a in range(A):
        O, d = _pixel_rays(geometry, pixel_size, U, V, a)
        for c, sg, amp in zip(np.asarray(centres, np.float64), sigmas, amplitudes):
            w = c - O
            along = d @ w
            dist2 = np.dot(w, w) - along ** 2
            out[a] += amp * sg * math.sqrt(2 * math.pi) * np.exp(-np.maximum(dist2, 0.0) / (2 * sg * sg))
    return out


def point_detector_coordinates(points, geometry, pixel_size):
    """Detector index coordinates (u_idx, v_idx) [A,n] and FDK weights [A,n] of world points,
    by intersecting the source->point line with the detector plane (FP64 vectors).
    The weight is (source-to-isocentre / source-to-point distance along the detector normal)^2."""
    ang, sid, sod, u0, v0, sx, sy, sz = [np.asarray(x, np.float64).reshape(-1) for x in
                                         (geometry if not isinstance(geometry, dict)
                                          else [geometry[k] for k in GEOMETRY_NAMES])]
    P = np.asarray(points, np.float64).reshape(-1, 3)
    A = ang.size
    ui = np.zeros((A, len(P)))
    vi = np.zeros((A, len(P)))
    w = np.zeros((A, len(P)))
    for a in range(A):
        c, s = math.cos(ang[a]), math.sin(ang[a])
        eu, en = np.array([c, s, 0.0]), np.array([-s, c, 0.0])
        S = sod[a] * en + np.array([sx[a], sy[a], sz[a]])
        plane = sod[a] - sid[a]                      # detector plane: x . en = plane
        lam = (plane - S @ en) / ((P - S) @ en)
        X = S + lam[:, None] * (P - S)
        ui[a] = (X @ eu) / pixel_size + u0[a]
        vi[a] = X[:, 2] / pixel_size + v0[a]
        w[a] = ((S @ en) / ((S - P) @ en)) ** 2
    return ui, vi, w


def preprocess(raw, dark=None, flat=None, layout="AVU", min_transmission=1e-6, log=True):
    """FP64 restatement of the detector preprocessing: [A,V,U] (or [A,U,V]) raw frames ->
    [A,U,V] line integrals -log(max((raw - dark) / (flat - dark), t_min)); flat <= dark
    gives -log(t_min).  The caller's step in the reference (README.md:97,108)."""
    x = np.asarray(raw, dtype=np.float64)
    if log:
        d = np.zeros(x.shape[1:]) if dark is None else np.asarray(dark, np.float64)
        f = np.ones(x.shape[1:]) if flat is None else np.asarray(flat, np.float64)
        den = f - d
        with np.errstate(divide="ignore", invalid="ignore"):
            t = np.where(den > 0, (x - d) / np.where(den > 0, den, 1.0), min_transmission)
        x = -np.log(np.maximum(t, min_transmission))
    return np.ascontiguousarray(x.transpose(0, 2, 1)) if layout == "AVU" else x

// Shared device helpers for the cone-beam CT kernels (gfx950).
//
// All files are compiled with -ffp-contract=off: every fused multiply-add is
// written out as fmaf(), so the float operation order below is fixed.  Ray
// set-up, sample positions, cells and fractions follow the FP32 reference
// schedule bit for bit; the trilinear interpolant is evaluated in a factored
// (nested-lerp) form that differs from the reference's weighted sum by FP32
// rounding only.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>

#include "cbct_hip.h"

namespace cbct {

// ---------------------------------------------------------------------------
// error state (thread-local message, integer status)
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_status(hipError_t err, const char* what);

#define CBCT_CHECK_ARG(cond, msg) \
    do { if (!(cond)) return ::cbct::fail(CBCT_EINVAL, (msg)); } while (0)
#define CBCT_HIP(call, what) \
    do { hipError_t e_ = (call); if (e_ != hipSuccess) return ::cbct::hip_status(e_, (what)); } while (0)
#define CBCT_LAUNCHED(what) \
    do { hipError_t e_ = hipGetLastError(); if (e_ != hipSuccess) return ::cbct::hip_status(e_, (what)); } while (0)

// ---------------------------------------------------------------------------
// per-view frame
// ---------------------------------------------------------------------------
// Everything a view contributes to a ray or a voxel.  Trig is evaluated in
// FP64 and rounded once; every other entry is the float expression of the
// reference (csrc/src_cbct_forwardprojector_cuda.cu:133-151,
// csrc/src_cbct_backprojector_cuda.cu:44-56).
struct alignas(16) ViewFrame {
    float c, s;          // cos / sin of the view angle
    float ox, oy, oz;    // source position  (-sod*s + sx, sod*c + sy, sz)
    float ns, nc;        // (sod - sid) * s, (sod - sid) * c
    float u0, v0;        // detector centre indices
    float sod, sid;
    float dsu, dsn, dsv; // source offset in the rotating frame (backprojector)
    float n;             // sod - sid
    float bad;           // 1: degenerate view (see guard_frame), every other field 0
};
static_assert(sizeof(ViewFrame) == 64, "ViewFrame is 16 floats");

struct GeomPtrs {
    const float* ang; const float* sid; const float* sod;
    const float* u0; const float* v0;
    const float* sx; const float* sy; const float* sz;
};

__device__ __forceinline__ ViewFrame make_frame(const GeomPtrs& g, int a) {
    ViewFrame f;
    const double beta = (double)g.ang[a];
    f.c = (float)cos(beta);
    f.s = (float)sin(beta);
    const float sod = g.sod[a], sid = g.sid[a], sx = g.sx[a], sy = g.sy[a], sz = g.sz[a];
    f.sod = sod;
    f.sid = sid;
    f.n = sod - sid;
    f.ox = fmaf(-sod, f.s, sx);
    f.oy = fmaf(sod, f.c, sy);
    f.oz = sz;
    f.ns = f.n * f.s;
    f.nc = f.n * f.c;
    f.u0 = g.u0[a];
    f.v0 = g.v0[a];
    f.dsu = fmaf(sx, f.c, sy * f.s);
    f.dsn = fmaf(-sx, f.s, sy * f.c);
    f.dsv = sz;
    f.bad = 0.f;
    return f;
}

// Which views the projector can march.  A view is degenerate -- zero projections and zero
// gradients, as the reference does for a degenerate ray
// (csrc/src_cbct_forwardprojector_cuda.cu:159-161) -- unless
//   (1) its eight geometry values and the frame made from them are finite;
//   (2) the unnormalised directions of its detector's corner rays have every component below
//       1e18 in magnitude.  rx and ry depend on u only, rz on v only, each through correctly
//       rounded operations with fixed-sign coefficients, so each is monotone in the pixel
//       index and every ray of the detector is bounded by the corners: its squared length is
//       finite and so is its unit direction;
//   (3) step >= 2^-22 * Tb with Tb = 2 (|O - C| + R) + 1, where O is the source, C the volume
//       box's centre and R its half-diagonal.  Every t of a sample (0 <= t <= t_max, and t_max
//       of a ray that meets the box is below Tb even with the slab test's rounding) is then
//       below Tb, so fl(t + step) >= t + step/2: each march ends after at most
//       2 (t_max - t_min) / step + 1 <= 2^23 + 1 samples.  (The reference's loop never ends when
//       step is below the spacing of floats at t, or when a NaN direction gives [0, FLT_MAX].)
// A degenerate view's frame is all zeros with bad = 1, so no later arithmetic meets a NaN or
// an infinity.  The tests must see NaNs: call guard_frame only from code compiled with NaN
// semantics (not under -fno-honor-nans).
struct FrameGuard {
    int on;                // 0: plain frames (the backprojector, whose loops are bounded by A)
    int U, V;
    float pix, step;
    float cx, cy, cz, R;   // volume box centre and half-diagonal, world units
};

inline FrameGuard make_guard_host(int64_t D, int64_t H, int64_t W, float vs, float ox, float oy, float oz,
                                  int64_t U, int64_t V, float pix, float step) {
    FrameGuard G;
    G.on = 1;
    G.U = (int)U; G.V = (int)V;
    G.pix = pix; G.step = step;
    // box [(-N/2 - 1/2) vs + off, (N/2 - 1/2) vs + off] per axis
    G.cx = (float)((double)ox - 0.5 * (double)vs);
    G.cy = (float)((double)oy - 0.5 * (double)vs);
    G.cz = (float)((double)oz - 0.5 * (double)vs);
    G.R = (float)(0.5 * (double)vs * std::sqrt((double)W * W + (double)H * H + (double)D * D));
    return G;
}

__device__ inline ViewFrame guard_frame(const GeomPtrs& g, int a, const FrameGuard& G) {
    ViewFrame f = make_frame(g, a);
    if (!G.on) return f;
    bool ok = isfinite(g.ang[a]) && isfinite(g.sid[a]) && isfinite(g.sod[a]) && isfinite(g.u0[a]) &&
              isfinite(g.v0[a]) && isfinite(g.sx[a]) && isfinite(g.sy[a]) && isfinite(g.sz[a]);
    ok = ok && isfinite(f.c) && isfinite(f.s) && isfinite(f.ox) && isfinite(f.oy) && isfinite(f.oz) &&
         isfinite(f.ns) && isfinite(f.nc) && isfinite(f.n) && isfinite(f.dsu) && isfinite(f.dsn);
    const float lim = 1e18f;
    for (int i = 0; i < 2; ++i) {
        // the expressions of make_ray at u = 0, U-1 and v = 0, V-1
        const float up = ((i ? (float)(G.U - 1) : 0.f) - f.u0) * G.pix;
        const float vp = ((i ? (float)(G.V - 1) : 0.f) - f.v0) * G.pix;
        const float rx = fmaf(up, f.c, -f.ns) - f.ox;
        const float ry = fmaf(up, f.s, f.nc) - f.oy;
        const float rz = vp - f.oz;
        ok = ok && fabsf(rx) <= lim && fabsf(ry) <= lim && fabsf(rz) <= lim;
    }
    if (ok) {
        const double dx = (double)f.ox - (double)G.cx, dy = (double)f.oy - (double)G.cy;
        const double dz = (double)f.oz - (double)G.cz;
        const double Tb = 2.0 * (sqrt(dx * dx + dy * dy + dz * dz) + (double)G.R) + 1.0;
        ok = (double)G.step >= Tb * 0x1p-22;
    }
    if (!ok) {
        f = ViewFrame{};
        f.bad = 1.f;
    }
    return f;
}

// Axis-aligned volume box and voxel-coordinate map (x = W axis, y = H, z = D).
// lo/hi: csrc/src_cbct_forwardprojector_cuda.cu:303-312.
struct Box {
    float lox, loy, loz, hix, hiy, hiz;
    float offx, offy, offz;
    float hx, hy, hz;   // W/2, H/2, D/2
    int W, H, D;
};

inline Box make_box_host(int64_t D, int64_t H, int64_t W, float vs, float ox, float oy, float oz) {
    Box b;
    b.W = (int)W; b.H = (int)H; b.D = (int)D;
    b.hx = (float)W * 0.5f; b.hy = (float)H * 0.5f; b.hz = (float)D * 0.5f;
    b.lox = std::fmaf((float)(-W) * 0.5f - 0.5f, vs, ox);
    b.loy = std::fmaf((float)(-H) * 0.5f - 0.5f, vs, oy);
    b.loz = std::fmaf((float)(-D) * 0.5f - 0.5f, vs, oz);
    b.hix = std::fmaf(b.hx - 0.5f, vs, ox);
    b.hiy = std::fmaf(b.hy - 0.5f, vs, oy);
    b.hiz = std::fmaf(b.hz - 0.5f, vs, oz);
    b.offx = ox; b.offy = oy; b.offz = oz;
    return b;
}

struct Ray {
    float dx, dy, dz, len, up;
};

// Unit direction from the source through detector pixel (u,v); false for a
// degenerate ray (csrc/src_cbct_forwardprojector_cuda.cu:138-169).
__device__ __forceinline__ bool make_ray(const ViewFrame& f, int u, int v, float pix, Ray& r) {
    if (f.bad != 0.f) return false;        // degenerate view (guard_frame)
    const float up = ((float)u - f.u0) * pix;
    const float vp = ((float)v - f.v0) * pix;
    const float tx = fmaf(up, f.c, -f.ns);
    const float ty = fmaf(up, f.s, f.nc);
    const float rx = tx - f.ox, ry = ty - f.oy, rz = vp - f.oz;
    const float len = sqrtf(fmaf(rz, rz, fmaf(ry, ry, rx * rx)));
    r.up = up;
    r.len = len;
    if (len <= 1e-8f) return false;
    const float inv = 1.0f / len;
    r.dx = rx * inv; r.dy = ry * inv; r.dz = rz * inv;
    return true;
}

// Slab update, csrc/src_cbct_forwardprojector_cuda.cu:55-83.
__device__ __forceinline__ bool slab(float o, float d, float lo, float hi, int axis,
                                     float& tmin, float& tmax, int& entry) {
    if (fabsf(d) < 1e-12f) return o >= lo && o <= hi;
    const float inv = 1.0f / d;
    float t1 = (lo - o) * inv, t2 = (hi - o) * inv;
    if (t1 > t2) { const float t = t1; t1 = t2; t2 = t; }
    if (t1 > tmin) { tmin = t1; entry = axis; }
    tmax = fminf(tmax, t2);
    return tmax >= tmin;
}

// csrc/src_cbct_forwardprojector_cuda.cu:86-112.
__device__ __forceinline__ bool intersect(const Box& b, const ViewFrame& f, const Ray& r,
                                          float& tmin, float& tmax, int& entry) {
    tmin = 0.0f; tmax = FLT_MAX; entry = -1;
    if (!slab(f.ox, r.dx, b.lox, b.hix, 0, tmin, tmax, entry)) return false;
    if (!slab(f.oy, r.dy, b.loy, b.hiy, 1, tmin, tmax, entry)) return false;
    if (!slab(f.oz, r.dz, b.loz, b.hiz, 2, tmin, tmax, entry)) return false;
    return tmax >= tmin;
}

// Voxel coordinate of a world coordinate.  POW2: voxel_size is a power of two,
// so the division of the forward kernel (:326-328) equals the reciprocal
// multiply of the adjoint (:425-427) and both are one fused operation.
template <bool DIV>
__device__ __forceinline__ float to_voxel(float p, float off, float vs, float inv_vs, float half) {
    if (DIV) return (p - off) / vs + half;
    return fmaf(p - off, inv_vs, half);
}

// floor(x) as an int, and x - floor(x) in one instruction (v_fract_f32; identical
// to the subtraction for every coordinate except -2^-25 < x < 0, where it gives
// the largest float below 1 instead of 1 -- a cell whose two x corners are the
// same clamped voxel, so nothing changes).
__device__ __forceinline__ int floor_int(float x) { return (int)floorf(x); }
__device__ __forceinline__ float frac(float x) { return __builtin_amdgcn_fractf(x); }

// Trilinear corner set of a sample: clamp-to-edge corner offsets and raw
// fractions (csrc/src_cbct_forwardprojector_cuda.cu:186-229).
struct Corners {
    int64_t i000;              // base offset (64-bit: a volume may exceed 2^31 voxels)
    int dx1, dy1, dz1;         // clamped +1 steps per axis (a plane H*W stays below 2^31)
    float fx, fy, fz;
};

__device__ __forceinline__ Corners corners(const Box& b, float x, float y, float z) {
    const int xr = floor_int(x), yr = floor_int(y), zr = floor_int(z);
    Corners c;
    c.fx = frac(x); c.fy = frac(y); c.fz = frac(z);
    const int x0 = max(0, min(xr, b.W - 1)), x1 = max(0, min(xr + 1, b.W - 1));
    const int y0 = max(0, min(yr, b.H - 1)), y1 = max(0, min(yr + 1, b.H - 1));
    const int z0 = max(0, min(zr, b.D - 1)), z1 = max(0, min(zr + 1, b.D - 1));
    const int P = b.H * b.W;
    c.i000 = (int64_t)z0 * P + (y0 * b.W + x0);
    c.dx1 = x1 - x0;
    c.dy1 = (y1 - y0) * b.W;
    c.dz1 = (z1 - z0) * P;
    return c;
}

// The same cell as unscaled indices: lower corner (clamped) and 0/1 steps.
struct Cell {
    int x0, y0, z0;
    int sx, sy, sz;            // 1 when the +1 corner is a distinct voxel, 0 at a clamped edge
    float fx, fy, fz;
};

__device__ __forceinline__ Cell cell_of(const Box& b, float x, float y, float z) {
    const int xr = floor_int(x), yr = floor_int(y), zr = floor_int(z);
    Cell c;
    c.fx = frac(x); c.fy = frac(y); c.fz = frac(z);
    c.x0 = max(0, min(xr, b.W - 1));
    c.y0 = max(0, min(yr, b.H - 1));
    c.z0 = max(0, min(zr, b.D - 1));
    c.sx = max(0, min(xr + 1, b.W - 1)) - c.x0;
    c.sy = max(0, min(yr + 1, b.H - 1)) - c.y0;
    c.sz = max(0, min(zr + 1, b.D - 1)) - c.z0;
    return c;
}

struct Taps { float v000, v100, v010, v110, v001, v101, v011, v111; };

// Eight taps at base offset i000 with strides (1, ry, rz) scaled by the 0/1 steps.
__device__ __forceinline__ Taps load_taps_strided(const float* __restrict__ base, int64_t i000,
                                                  int dx, int dy, int dz) {
    const float* p = base + i000;
    Taps t;
    t.v000 = p[0];           t.v100 = p[dx];
    t.v010 = p[dy];          t.v110 = p[dy + dx];
    t.v001 = p[dz];          t.v101 = p[dz + dx];
    t.v011 = p[dz + dy];     t.v111 = p[dz + dy + dx];
    return t;
}

__device__ __forceinline__ Taps load_taps(const float* __restrict__ vol, const Corners& c) {
    const float* p = vol + c.i000;
    Taps t;
    t.v000 = p[0];               t.v100 = p[c.dx1];
    t.v010 = p[c.dy1];           t.v110 = p[c.dy1 + c.dx1];
    t.v001 = p[c.dz1];           t.v101 = p[c.dz1 + c.dx1];
    t.v011 = p[c.dz1 + c.dy1];   t.v111 = p[c.dz1 + c.dy1 + c.dx1];
    return t;
}

struct Weights { float w000, w100, w010, w110, w001, w101, w011, w111; };

__device__ __forceinline__ Weights tri_weights(float fx, float fy, float fz) {
    const float mx = 1.0f - fx, my = 1.0f - fy, mz = 1.0f - fz;
    Weights w;
    w.w000 = mx * my * mz; w.w100 = fx * my * mz; w.w010 = mx * fy * mz; w.w110 = fx * fy * mz;
    w.w001 = mx * my * fz; w.w101 = fx * my * fz; w.w011 = mx * fy * fz; w.w111 = fx * fy * fz;
    return w;
}

// Trilinear value and spatial gradient (voxel units) of the eight taps, as
// nested linear interpolations along x, then y, then z.  This is the
// reference's interpolant (csrc/src_cbct_forwardprojector_cuda.cu:186-229) and
// its analytic derivative (:242-250) in factored form: 14 operations for the
// value and 8 more for the gradient (the weighted corner sum takes 23 and 36);
// the results differ from the weighted sum by FP32 rounding only.
__device__ __forceinline__ void tri_eval(const Taps& t, float fx, float fy, float fz,
                                         float& val, float& gx, float& gy, float& gz) {
    const float d00 = t.v100 - t.v000, d10 = t.v110 - t.v010;      // x differences per (y, z)
    const float d01 = t.v101 - t.v001, d11 = t.v111 - t.v011;
    const float c00 = fmaf(fx, d00, t.v000), c10 = fmaf(fx, d10, t.v010);
    const float c01 = fmaf(fx, d01, t.v001), c11 = fmaf(fx, d11, t.v011);
    const float e0d = c10 - c00, e1d = c11 - c01;                    // y differences per z
    const float e0 = fmaf(fy, e0d, c00), e1 = fmaf(fy, e1d, c01);
    gz = e1 - e0;
    val = fmaf(fz, gz, e0);
    gy = fmaf(fz, e1d - e0d, e0d);
    const float x0 = fmaf(fy, d10 - d00, d00), x1 = fmaf(fy, d11 - d01, d01);
    gx = fmaf(fz, x1 - x0, x0);
}

__device__ __forceinline__ float tri_value(const Taps& t, float fx, float fy, float fz) {
    float val, gx, gy, gz;
    tri_eval(t, fx, fy, fz, val, gx, gy, gz);
    return val;
}

// Wave-level (64 lanes) sum.
__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m, 64);
    return x;
}

// Is x a power of two (so that division by it is exact scaling)?
inline bool is_pow2(float x) {
    int e = 0;
    const float m = std::frexp(x, &e);
    return m == 0.5f;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Checks shared by all entry points.
int check_geometry(const float* const* g, int64_t A);
int check_positive_finite(float x, const char* name);
int check_finite(float x, const char* name);
// step_size against the FP32 resolution of the ray parameter over the volume: a step below
// 2^-22 (volume diagonal + 1) leaves every view degenerate (guard_frame), so it is refused.
int check_step_resolution(float step, const FrameGuard& G);
// The device of the calling thread, after checking that the stream (if not the null stream) belongs
// to it: the entry points that keep per-device state (the allocation-free adjoint's scratch, the
// calibration's captured graphs) look that state up on the current device.
int current_stream_device(hipStream_t stream, int* device);

__global__ void frames_kernel(GeomPtrs g, int A, FrameGuard guard, ViewFrame* __restrict__ out);
__global__ void fill_kernel(float* __restrict__ p, int64_t n, float value);
// Frames of the A views; with a guard (projector) degenerate views are flagged (guard_frame).
int launch_frames(const GeomPtrs& g, int64_t A, ViewFrame* out, hipStream_t stream,
                  const FrameGuard* guard = nullptr);
int launch_fill(float* p, int64_t n, float value, hipStream_t stream);
// Kernels of cbct_fp.hip shared with the backprojector backward: zero n int64 words; max |x|
// over n floats into *maxbits (as float bits, *maxbits cleared beforehand); ggeom[k][a] = the
// fixed-order sum of partials[(a * tiles + t) * 8 + k] over t.
int launch_zero_u64(unsigned long long* p, int64_t n, hipStream_t stream);
int launch_absmax(const float* x, int64_t n, unsigned* maxbits, hipStream_t stream);
int launch_reduce_geometry(const float* partials, int64_t A, int64_t tiles, float* ggeom, hipStream_t stream);
// The deterministic setting of a call: the caller's options, or the process-wide default
// (cbct_fp.hip).
int call_deterministic(const cbct_fp_options* opt, int* det);

}  // namespace cbct

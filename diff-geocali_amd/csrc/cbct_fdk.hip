// FDK preweight + Hann-windowed ramp filter, and the chunked FDK
// reconstruction, for gfx950.
//
// The reference filters through a zero-padded FFT of length
// pad = 1 << bit_length(2U-1) (cbct_backprojector/filtering.py:87-104).  Since
// pad >= 2U-1, that circular convolution restricted to the first U outputs is
// exactly a linear convolution along U with the kernel
//     h_eff[d] = (1/pad) * sum_k R[k] * Hann[k] * cos(2 pi k d / pad),  |d| < U,
//     R[k]     = sum_n h[n] cos(2 pi k n / pad)   (real part of FFT of the ramp)
// so the filter here is a direct convolution with 2U-1 taps computed on the
// device (FP32, compensated sums), applied in LDS-staged register tiles.
#include "cbct_common.h"

#include <algorithm>

namespace cbct {

static int64_t fdk_pad(int64_t U) {
    int64_t n = 2 * U - 1, bits = 0;
    while (n) { ++bits; n >>= 1; }
    return int64_t(1) << bits;
}

// Spatial ramp kernel of the reference (filtering.py:89-94), wrapped index n.
__device__ __forceinline__ float ramp_tap(int n, int pad, float pix) {
    const int nn = n > pad / 2 ? n - pad : n;
    if (nn == 0) return 1.0f / (4.0f * pix * pix);
    if (nn & 1) {
        const float d = 3.14159265358979323846f * (float)nn * pix;
        return -1.0f / (d * d);
    }
    return 0.0f;
}

// RH[k] = Re(FFT(ramp))[k] * Hann[k] (filtering.py:95-101).
__global__ __launch_bounds__(256) void ramp_response_kernel(int pad, float pix, float alpha,
                                                            float* __restrict__ rh) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= pad) return;
    float sum = 0.0f, comp = 0.0f;   // Kahan sum
    for (int n = 0; n < pad; ++n) {
        const int m = (int)(((int64_t)k * n) & (pad - 1));
        const float term = ramp_tap(n, pad, pix) * cospif(2.0f * (float)m / (float)pad);
        const float y = term - comp;
        const float t = sum + y;
        comp = (t - sum) - y;
        sum = t;
    }
    const float f = (float)(k < (pad + 1) / 2 ? k : k - pad) / (float)pad;   // fftfreq
    const float hann = fabsf(f) <= alpha * 0.5f ? 0.5f + 0.5f * cospif(2.0f * f / alpha) : 0.0f;
    rh[k] = sum * hann;
}

// taps[j], j in [0, 2U-1): h_eff[j - (U-1)], placed at offset TAP_PAD of a
// zero-padded array of length 2U-1 + 2*TAP_PAD.
constexpr int TAP_PAD = 64;

__global__ __launch_bounds__(256) void taps_kernel(const float* __restrict__ rh, int pad, int U,
                                                   float* __restrict__ taps) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int len = 2 * U - 1 + 2 * TAP_PAD;
    if (j >= len) return;
    const int d = j - TAP_PAD - (U - 1);
    if (d <= -U || d >= U) { taps[j] = 0.0f; return; }
    const int ad = d < 0 ? -d : d;
    float sum = 0.0f, comp = 0.0f;
    for (int k = 0; k < pad; ++k) {
        const int m = (int)(((int64_t)k * ad) & (pad - 1));
        const float term = rh[k] * cospif(2.0f * (float)m / (float)pad);
        const float y = term - comp;
        const float t = sum + y;
        comp = (t - sum) - y;
        sum = t;
    }
    taps[j] = sum / (float)pad;
}

// Filter tile: 64 outputs along u x 64 along v of one view; wave w owns
// u in [ub + 16w, ub + 16w + 16), lane owns one v.  Inputs are staged 64 u at a
// time in LDS with the cosine preweight applied; each 16-input sub-block uses
// a 31-tap register window, so one uniform LDS read feeds 16 FMAs.
constexpr int FT_U = 64, FT_V = 64;

__global__ __launch_bounds__(256) void fdk_filter_kernel(
    const float* __restrict__ proj, int A, int U, int V, float pix,
    const float* __restrict__ sod, const float* __restrict__ sid,
    const float* __restrict__ u0a, const float* __restrict__ v0a, const float* __restrict__ pw,
    const float* __restrict__ taps_g, float* __restrict__ out) {
    extern __shared__ float smem[];
    const int ntaps = 2 * U - 1 + 2 * TAP_PAD;
    float* taps = smem;                           // ntaps
    float* X = smem + ((ntaps + 3) & ~3);         // [FT_U][FT_V + 1]
    const int tiles_v = (V + FT_V - 1) / FT_V;
    const int tiles_u = (U + FT_U - 1) / FT_U;
    const int a = blockIdx.x / (tiles_u * tiles_v);
    const int r = blockIdx.x % (tiles_u * tiles_v);
    const int ub = (r / tiles_v) * FT_U;
    const int vb = (r % tiles_v) * FT_V;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < ntaps; i += 256) taps[i] = taps_g[i];
    const float sid_a = sid[a], sod_a = sod[a];
    // detector centre of the cosine preweight: the geometric centre (the reference's
    // convention), or this view's det_u0/det_v0 when given
    const float cu = u0a ? u0a[a] : (float)(U - 1) * 0.5f, cv = v0a ? v0a[a] : (float)(V - 1) * 0.5f;
    const float sid2 = sid_a * sid_a;
    const int v = vb + lane;
    float acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
    const int u0 = ub + 16 * wave;               // first output u of this wave
    const float* pa = proj + (int64_t)a * U * V;
    for (int c = 0; c < U; c += FT_U) {
        __syncthreads();
        // stage X[uu][vv] = preweighted proj[a, c+uu, vb+vv]
        for (int i = threadIdx.x; i < FT_U * FT_V; i += 256) {
            const int uu = i / FT_V, vv = i % FT_V;
            const int gu = c + uu, gv = vb + vv;
            float x = 0.0f;
            if (gu < U && gv < V) {
                const float up = ((float)gu - cu) * pix;
                const float vp = ((float)gv - cv) * pix;
                const float dist = sqrtf((sid2 + up * up) + vp * vp);
                x = pa[(int64_t)gu * V + gv] * (sid_a / dist);
                if (pw) x *= pw[(int64_t)a * U + gu];
            }
            X[uu * (FT_V + 1) + vv] = x;
        }
        __syncthreads();
#pragma unroll 1
        for (int s = 0; s < FT_U; s += 16) {
            // window w[i] = h_eff[(u0 - (c + s)) + (i - 15)], i in [0, 31)
            const int base = TAP_PAD + (U - 1) + (u0 - (c + s)) - 15;
            float w[31];
#pragma unroll
            for (int i = 0; i < 31; ++i) w[i] = taps[base + i];
#pragma unroll
            for (int rr = 0; rr < 16; ++rr) {
                const float xv = X[(s + rr) * (FT_V + 1) + lane];
#pragma unroll
                for (int j = 0; j < 16; ++j) acc[j] = fmaf(w[15 + j - rr], xv, acc[j]);
            }
        }
    }
    if (v < V) {
        const float ratio = sid_a / sod_a;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int u = u0 + j;
            if (u < U) out[((int64_t)a * U + u) * V + v] = (acc[j] * pix) * ratio;
        }
    }
}

// Parker short-scan weights per (view, detector column) (Parker 1982; Kak & Slaney ch. 3):
// with beta = angle - angle_0, fan angle gamma = atan(u_p / sid) (u_p along the detector u
// axis from the preweight centre) and delta = (range - pi) / 2 for the scan range
// angle_{A-1} - angle_0, the ray (beta, gamma) is measured again as (beta + pi + 2 gamma, -gamma)
// in this geometry, and
//   w = sin^2(pi/4 * beta / (delta - gamma))              for 0 <= beta < 2 delta - 2 gamma
//     = 1                                                  up to pi - 2 gamma
//     = sin^2(pi/4 * (pi + 2 delta - beta) / (delta + gamma)) up to pi + 2 delta
// makes every line's weights sum to one.  view_scale (reconstruction): the weights also carry
// the ratio of the short scan's view weight, the mean angular step (angle_{A-1} - angle_0)/(A-1),
// to the pi/A that the backprojection applies, so the angular range never has to be read back
// to the host.
__global__ void parker_weights_kernel(const float* __restrict__ angles, int A, int U,
                                      const float* __restrict__ u0a, float pix,
                                      const float* __restrict__ sid, int view_scale, float* __restrict__ w) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)A * U) return;
    const int a = (int)(i / U), u = (int)(i % U);
    const float pi = 3.14159265358979f;
    const float beta = angles[a] - angles[0];
    const float delta = fmaxf((angles[A - 1] - angles[0] - pi) * 0.5f, 1e-9f);
    const float cu = u0a ? u0a[a] : (float)(U - 1) * 0.5f;
    const float g = atanf(((float)u - cu) * pix / sid[a]);
    float val = 0.f;
    if (beta >= 0.f && beta < 2.f * delta - 2.f * g) {
        const float t = sinf(pi / 4.f * beta / (delta - g));
        val = t * t;
    } else if (beta >= 2.f * delta - 2.f * g && beta <= pi - 2.f * g) {
        val = 1.f;
    } else if (beta > pi - 2.f * g && beta <= pi + 2.f * delta) {
        const float t = sinf(pi / 4.f * (pi + 2.f * delta - beta) / (delta + g));
        val = t * t;
    }
    if (view_scale) {
        const double step = ((double)angles[A - 1] - (double)angles[0]) / (double)(A - 1);
        val = (float)((double)val * step * ((double)A / 3.14159265358979323846));
    }
    w[i] = val;
}

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct FdkLayout {
    size_t frames, rh, taps, parker, filtered, total;
};

static FdkLayout fdk_layout(int64_t A, int64_t U, int64_t V) {
    FdkLayout L;
    L.frames = 0;
    L.rh = align256((size_t)A * sizeof(ViewFrame));
    L.taps = L.rh + align256((size_t)fdk_pad(U) * sizeof(float));
    L.parker = L.taps + align256((size_t)(2 * U - 1 + 2 * TAP_PAD) * sizeof(float));
    L.filtered = L.parker + align256((size_t)A * U * sizeof(float));
    L.total = L.filtered + (size_t)A * U * V * sizeof(float);
    return L;
}

static int launch_filter(const float* proj, int64_t A, int64_t U, int64_t V, float pix,
                         const float* sod, const float* sid, float alpha, const float* u0a, const float* v0a,
                         const float* angles, float* out, float* rh, float* taps, float* pw,
                         hipStream_t stream, int view_scale = 0) {
    const int pad = (int)fdk_pad(U);
    ramp_response_kernel<<<(pad + 255) / 256, 256, 0, stream>>>(pad, pix, alpha, rh);
    CBCT_LAUNCHED("ramp response kernel launch");
    const int ntaps = (int)(2 * U - 1 + 2 * TAP_PAD);
    taps_kernel<<<(ntaps + 255) / 256, 256, 0, stream>>>(rh, pad, (int)U, taps);
    CBCT_LAUNCHED("filter taps kernel launch");
    if (angles) {
        parker_weights_kernel<<<(unsigned)ceil_div(A * U, (int64_t)256), 256, 0, stream>>>(angles, (int)A, (int)U,
                                                                                          u0a, pix, sid, view_scale,
                                                                                          pw);
        CBCT_LAUNCHED("Parker weights kernel launch");
    }
    const int64_t blocks = A * ceil_div(U, FT_U) * ceil_div(V, FT_V);
    const size_t shmem = (size_t)(((ntaps + 3) & ~3) + FT_U * (FT_V + 1)) * sizeof(float);
    fdk_filter_kernel<<<(unsigned)blocks, 256, shmem, stream>>>(proj, (int)A, (int)U, (int)V, pix,
                                                                sod, sid, u0a, v0a, angles ? pw : nullptr, taps,
                                                                out);
    CBCT_LAUNCHED("FDK filter kernel launch");
    return CBCT_OK;
}

static int check_fdk_common(int64_t A, int64_t U, int64_t V, float pix, float alpha) {
    CBCT_CHECK_ARG(A >= 1 && U >= 1 && V >= 1, "projections must have a nonempty [A, U, V] shape");
    CBCT_CHECK_ARG(A * U * V <= INT32_MAX, "projections exceed the 2^31-1 element limit");
    CBCT_CHECK_ARG(U <= 4096, "U exceeds the filter's 4096-pixel limit");
    int st;
    if ((st = check_positive_finite(pix, "pixel_size"))) return st;
    if ((st = check_positive_finite(alpha, "alpha"))) return st;
    CBCT_CHECK_ARG(alpha <= 1.0f, "alpha must satisfy 0 < alpha <= 1");
    return CBCT_OK;
}

}  // namespace cbct

using namespace cbct;

extern "C" size_t cbct_fdk_workspace_bytes(int64_t A, int64_t U, int64_t V) {
    if (A < 1 || U < 1 || V < 1) return 0;
    return fdk_layout(A, U, V).total;
}

extern "C" int cbct_fdk_filter_ex(const float* projections, int64_t A, int64_t U, int64_t V,
                                  float pixel_size, const float* sod, const float* sid, float alpha,
                                  const float* det_u0, const float* det_v0, const float* angles,
                                  float* filtered, void* workspace, size_t workspace_bytes, void* stream_) {
    int st;
    if ((st = check_fdk_common(A, U, V, pixel_size, alpha))) return st;
    CBCT_CHECK_ARG(projections && sod && sid && filtered, "projection, distance and output pointers must be non-null");
    CBCT_CHECK_ARG((det_u0 == nullptr) == (det_v0 == nullptr), "det_u0 and det_v0 must both be given or both be null");
    CBCT_CHECK_ARG(!angles || A >= 2, "a short-scan (Parker) filter needs at least two views");
    const FdkLayout L = fdk_layout(A, U, V);
    CBCT_CHECK_ARG(workspace != nullptr && workspace_bytes >= L.filtered, "workspace too small for cbct_fdk_filter");
    char* ws = (char*)workspace;
    return launch_filter(projections, A, U, V, pixel_size, sod, sid, alpha, det_u0, det_v0, angles, filtered,
                         (float*)(ws + L.rh), (float*)(ws + L.taps), (float*)(ws + L.parker), (hipStream_t)stream_);
}

extern "C" int cbct_fdk_filter(const float* projections, int64_t A, int64_t U, int64_t V,
                               float pixel_size, const float* sod, const float* sid, float alpha,
                               float* filtered, void* workspace, size_t workspace_bytes, void* stream_) {
    return cbct_fdk_filter_ex(projections, A, U, V, pixel_size, sod, sid, alpha, nullptr, nullptr, nullptr,
                              filtered, workspace, workspace_bytes, stream_);
}

namespace cbct {
int launch_bp_forward_fdk(const float* proj, const ViewFrame* frames, int64_t A, int64_t U, int64_t V,
                          int64_t D, int64_t H, int64_t W, float pix, float vs, float ox, float oy,
                          float oz, int chunk, float scale, float* vol, hipStream_t stream);
}

extern "C" int cbct_fdk_reconstruct_ex(const float* projections, int64_t A, int64_t U, int64_t V,
                                       const float* angles, const float* sid, const float* sod,
                                       const float* det_u0, const float* det_v0,
                                       const float* src_x, const float* src_y, const float* src_z,
                                       float pixel_size, int64_t D, int64_t H, int64_t W, float voxel_size,
                                       float off_x, float off_y, float off_z,
                                       float alpha, int64_t angle_chunk_size, int short_scan, int offset_preweight,
                                       float* volume, void* workspace, size_t workspace_bytes, void* stream_) {
    int st;
    if ((st = check_fdk_common(A, U, V, pixel_size, alpha))) return st;
    CBCT_CHECK_ARG(U >= 2 && V >= 2, "reconstruct_fdk requires detector dimensions U and V >= 2");
    const float* gg[8] = {angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z};
    if ((st = check_geometry(gg, A))) return st;
    CBCT_CHECK_ARG(projections && volume, "projections and volume must be non-null");
    CBCT_CHECK_ARG(angle_chunk_size >= 1, "angle_chunk_size must be at least 1");
    CBCT_CHECK_ARG(D > 0 && H > 0 && W > 0 && D * H * W <= INT32_MAX,
                   "volume_shape must be positive and within the 2^31-1 element limit");
    CBCT_CHECK_ARG(D <= 64 * 65535, "D exceeds the backprojector's launch limit of 4194240 voxels");
    CBCT_CHECK_ARG(!short_scan || A >= 2, "a short-scan reconstruction needs at least two views");
    if ((st = check_positive_finite(voxel_size, "voxel_size"))) return st;
    if ((st = check_finite(off_x, "off_x")) || (st = check_finite(off_y, "off_y")) ||
        (st = check_finite(off_z, "off_z")))
        return st;
    const FdkLayout L = fdk_layout(A, U, V);
    CBCT_CHECK_ARG(workspace != nullptr && workspace_bytes >= L.total, "workspace too small for cbct_fdk_reconstruct");
    hipStream_t stream = (hipStream_t)stream_;
    char* ws = (char*)workspace;
    ViewFrame* frames = (ViewFrame*)(ws + L.frames);
    float* filtered = (float*)(ws + L.filtered);
    // full scan: each view weighs (1/2) * 2 pi / A; a Parker short scan weighs the mean angular
    // step, which the Parker weights carry as a factor (parker_weights_kernel), so the range is
    // never read back to the host
    const double scale = 3.14159265358979323846 / (double)A;
    const GeomPtrs gp{angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z};
    if ((st = launch_frames(gp, A, frames, stream))) return st;
    if ((st = launch_filter(projections, A, U, V, pixel_size, sod, sid, alpha, offset_preweight ? det_u0 : nullptr,
                            offset_preweight ? det_v0 : nullptr, short_scan ? angles : nullptr, filtered,
                            (float*)(ws + L.rh), (float*)(ws + L.taps), (float*)(ws + L.parker), stream,
                            short_scan ? 1 : 0)))
        return st;
    const int chunk = (int)std::min<int64_t>(angle_chunk_size, A);
    return launch_bp_forward_fdk(filtered, frames, A, U, V, D, H, W, pixel_size, voxel_size,
                                 off_x, off_y, off_z, chunk, (float)scale, volume, stream);
}

extern "C" int cbct_fdk_reconstruct(const float* projections, int64_t A, int64_t U, int64_t V,
                                    const float* angles, const float* sid, const float* sod,
                                    const float* det_u0, const float* det_v0,
                                    const float* src_x, const float* src_y, const float* src_z,
                                    float pixel_size, int64_t D, int64_t H, int64_t W, float voxel_size,
                                    float off_x, float off_y, float off_z,
                                    float alpha, int64_t angle_chunk_size,
                                    float* volume, void* workspace, size_t workspace_bytes, void* stream_) {
    return cbct_fdk_reconstruct_ex(projections, A, U, V, angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z,
                                   pixel_size, D, H, W, voxel_size, off_x, off_y, off_z, alpha, angle_chunk_size,
                                   0, 0, volume, workspace, workspace_bytes, stream_);
}

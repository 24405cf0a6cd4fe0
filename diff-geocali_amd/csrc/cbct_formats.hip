// Detector data -> line integrals, the input format of the projector pair (gfx950).
//
// The reference expects projections [A,U,V] fp32 that are already dark/flat corrected and
// converted to line integrals, and asks callers with [A,V,U] data to transpose first
// (README.md:97,108; docs/USAGE.md:247-251).  cbct_preprocess does both in one HBM pass:
//   t = (raw - dark) / (flat - dark),  p = -log(max(t, t_min))     (or p = raw without log)
// with the [A,V,U] -> [A,U,V] transpose staged through LDS so that reads run along U and
// writes along V, both coalesced.  One workgroup = one 64x64 (V x U) tile of one view.
#include "cbct_common.h"

namespace cbct {
namespace {

constexpr int PT = 64;              // tile edge
constexpr int PT_PAD = PT + 1;      // LDS row stride (odd: transposed reads hit distinct banks)

// -log of the transmission of one pixel; a pixel with flat <= dark gives -log(t_min)
__device__ __forceinline__ float line_integral(float x, const float* dark, const float* flat, int64_t pix,
                                               float tmin, float log_tmin) {
    const float d = dark ? dark[pix] : 0.f;
    const float den = (flat ? flat[pix] : 1.f) - d;
    return den > 0.f ? -__logf(fmaxf((x - d) / den, tmin)) : -log_tmin;
}

template <typename T, int TRANSPOSE>
__global__ __launch_bounds__(256) void preprocess_kernel(const T* __restrict__ raw, const float* __restrict__ dark,
                                                         const float* __restrict__ flat, int64_t A, int V, int U,
                                                         float tmin, int take_log, float* __restrict__ out) {
    __shared__ float tile[PT * PT_PAD];
    const int tiles_u = (U + PT - 1) / PT;
    const int tu = blockIdx.x % tiles_u, tv = blockIdx.x / tiles_u;
    const int64_t a = blockIdx.y;
    const int u0 = tu * PT, v0 = tv * PT;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;      // 64 lanes along the fast axis, 4 rows
    const float log_tmin = __logf(tmin);
    const T* src = raw + a * V * U;
    const int f = u0 + tx;
    if (TRANSPOSE) {
#pragma unroll 4
        for (int r = ty; r < PT; r += 4) {
            const int s = v0 + r;
            float p = 0.f;
            if (s < V && f < U) {
                const int64_t pix = (int64_t)s * U + f;
                const float x = (float)src[pix];
                p = take_log ? line_integral(x, dark, flat, pix, tmin, log_tmin) : x;
            }
            tile[r * PT_PAD + tx] = p;
        }
        __syncthreads();
#pragma unroll 4
        for (int r = ty; r < PT; r += 4) {
            const int u = u0 + r, v = v0 + tx;
            if (u < U && v < V) out[(a * U + u) * V + v] = tile[tx * PT_PAD + r];
        }
    } else {
#pragma unroll 4
        for (int r = ty; r < PT; r += 4) {
            const int s = v0 + r;
            if (s < V && f < U) {
                const int64_t pix = (int64_t)s * U + f;
                const float x = (float)src[pix];
                out[a * V * U + pix] = take_log ? line_integral(x, dark, flat, pix, tmin, log_tmin) : x;
            }
        }
    }
}

// Transposing variant with paired accesses on the same 64x64 tile: each lane reads two
// consecutive u (one 4-byte uint16 pair or 8-byte float pair) and writes two consecutive v
// (one 8-byte store), half the memory instructions of the element-wise kernel.  Needs an even
// number of rows and columns and aligned buffers.
constexpr int PW = 64;              // tile width along u
constexpr int PW_PAD = PW + 1;

template <typename T> struct Pair;
template <> struct Pair<uint16_t> { typedef ushort2 type; };
template <> struct Pair<float> { typedef float2 type; };

template <typename T>
__global__ __launch_bounds__(256) void preprocess_pair_kernel(const T* __restrict__ raw, const float* __restrict__ dark,
                                                              const float* __restrict__ flat, int V, int U,
                                                              float tmin, int take_log, float* __restrict__ out) {
    __shared__ float tile[PT * PW_PAD];
    const int tiles_u = (U + PW - 1) / PW;
    const int tu = blockIdx.x % tiles_u, tv = blockIdx.x / tiles_u;
    const int64_t a = blockIdx.y;
    const int u0 = tu * PW, v0 = tv * PT;
    const float log_tmin = __logf(tmin);
    const T* src = raw + a * (int64_t)V * U;
    {
        const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;     // 32 lanes along u pairs, 8 v rows
        const int u = u0 + 2 * tx;
#pragma unroll 4
        for (int r = ty; r < PT; r += 8) {
            const int v = v0 + r;
            float p0 = 0.f, p1 = 0.f;
            if (v < V && u < U) {
                const int64_t pix = (int64_t)v * U + u;
                const typename Pair<T>::type x = *reinterpret_cast<const typename Pair<T>::type*>(src + pix);
                p0 = (float)x.x;
                p1 = (float)x.y;
                if (take_log) {
                    p0 = line_integral(p0, dark, flat, pix, tmin, log_tmin);
                    p1 = line_integral(p1, dark, flat, pix + 1, tmin, log_tmin);
                }
            }
            tile[r * PW_PAD + 2 * tx] = p0;
            tile[r * PW_PAD + 2 * tx + 1] = p1;
        }
    }
    __syncthreads();
    {
        const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;     // 32 lanes along v pairs, 8 u rows
        const int v = v0 + 2 * tx;
#pragma unroll 4
        for (int r = ty; r < PW; r += 8) {
            const int u = u0 + r;
            if (u < U && v < V) {
                const float2 val = make_float2(tile[(2 * tx) * PW_PAD + r], tile[(2 * tx + 1) * PW_PAD + r]);
                *reinterpret_cast<float2*>(out + (a * U + u) * (int64_t)V + v) = val;
            }
        }
    }
}

// Transposing variant for frames whose rows hold a multiple of 8 pixels and whose columns a
// multiple of 4: each lane reads 8 consecutive u (16 bytes of uint16) and writes 4 consecutive v
// (16 bytes), and a workgroup runs NV views of one 64x64 tile, keeping its pixels' dark value and
// reciprocal flat-dark gain in registers (the fields are read once per NV views instead of once
// per view).
template <typename T> struct Oct;
template <> struct Oct<uint16_t> {
    static __device__ __forceinline__ void load(const uint16_t* p, float* x) {
        const uint4 w = *reinterpret_cast<const uint4*>(p);
        const unsigned q[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) { x[2 * k] = (float)(q[k] & 0xffffu); x[2 * k + 1] = (float)(q[k] >> 16); }
    }
};
template <> struct Oct<float> {
    static __device__ __forceinline__ void load(const float* p, float* x) {
        const float4 a = *reinterpret_cast<const float4*>(p), b = *reinterpret_cast<const float4*>(p + 4);
        x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w; x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
    }
};

template <typename T, int NV>
__global__ __launch_bounds__(256) void preprocess_views_kernel(const T* __restrict__ raw, const float* __restrict__ dark,
                                                               const float* __restrict__ flat, int A, int V, int U,
                                                               float tmin, int take_log, float* __restrict__ out) {
    __shared__ float tile[PT * PT_PAD];          // [u][v] of the current view
    const int tiles_u = (U + PT - 1) / PT;
    const int tu = blockIdx.x % tiles_u, tv = blockIdx.x / tiles_u;
    const int u0 = tu * PT, v0 = tv * PT;
    const int a0 = blockIdx.y * NV;
    const int t = threadIdx.x;
    const int ru = (t & 7) * 8, rv = t >> 3;     // reads: u0+ru .. +7 of rows v0+rv and v0+rv+32
    const int wv = (t & 15) * 4, wu = t >> 4;    // writes: v0+wv .. +3 of u-rows u0+wu+16h
    float dd[2][8], gi[2][8];
    bool rok[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int v = v0 + rv + 32 * h, u = u0 + ru;
        rok[h] = v < V && u < U;                 // U % 8 == 0: a chunk is wholly inside or outside
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            dd[h][k] = 0.f;
            gi[h][k] = 1.f;
            if (rok[h] && take_log) {
                const int64_t pix = (int64_t)v * U + u + k;
                const float d = dark ? dark[pix] : 0.f;
                const float den = (flat ? flat[pix] : 1.f) - d;
                // a pixel with flat <= dark: t = 0 -> max(t, t_min) = t_min
                dd[h][k] = den > 0.f ? d : 0.f;
                gi[h][k] = den > 0.f ? 1.0f / den : 0.f;
            }
        }
    }
    // the next view's raw values are loaded while the current view is written out
    float xn[2][8];
    auto load_view = [&](int a) {
        const T* src = raw + (int64_t)a * V * U;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int k = 0; k < 8; ++k) xn[h][k] = 0.f;
            if (rok[h]) Oct<T>::load(src + (int64_t)(v0 + rv + 32 * h) * U + u0 + ru, xn[h]);
        }
    };
    if (a0 < A) load_view(a0);
    for (int j = 0; j < NV; ++j) {
        const int a = a0 + j;
        if (a >= A) break;                       // uniform over the workgroup
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const float x = xn[h][k];
                const float p = take_log ? -__logf(fmaxf((x - dd[h][k]) * gi[h][k], tmin)) : x;
                tile[(ru + k) * PT_PAD + rv + 32 * h] = p;
            }
        }
        __syncthreads();
        if (j + 1 < NV && a + 1 < A) load_view(a + 1);
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const int ul = wu + 16 * h, u = u0 + ul, v = v0 + wv;
            if (u < U && v < V) {                // V % 4 == 0: a chunk is wholly inside or outside
                const float* r = tile + ul * PT_PAD + wv;
                *reinterpret_cast<float4*>(out + ((int64_t)a * U + u) * V + v) = make_float4(r[0], r[1], r[2], r[3]);
            }
        }
        __syncthreads();
    }
}

}  // namespace
}  // namespace cbct

using namespace cbct;

extern "C" int cbct_preprocess(const void* raw, int raw_dtype, const float* dark, const float* flat, int64_t A,
                               int64_t rows, int64_t cols, int transpose, float min_transmission, int take_log,
                               float* out, void* stream) {
    CBCT_CHECK_ARG(raw && out, "preprocess needs input and output buffers");
    CBCT_CHECK_ARG(A > 0 && rows > 0 && cols > 0, "projection dimensions must be positive");
    CBCT_CHECK_ARG(A <= 65535 && rows * cols <= INT32_MAX && rows < (1 << 24) && cols < (1 << 24),
                   "projection dimensions exceed the launch limits");
    CBCT_CHECK_ARG(raw_dtype == 0 || raw_dtype == 1, "raw_dtype must be 0 (uint16) or 1 (float32)");
    CBCT_CHECK_ARG(transpose == 0 || transpose == 1, "transpose must be 0 or 1");
    CBCT_CHECK_ARG(min_transmission > 0.f && min_transmission <= 1.f, "min_transmission must be in (0, 1]");
    const int64_t tiles = ceil_div(rows, (int64_t)PT) * ceil_div(cols, (int64_t)PT);
    CBCT_CHECK_ARG(tiles <= INT32_MAX, "too many tiles");
    const dim3 grid((unsigned)tiles, (unsigned)A);
    hipStream_t s = (hipStream_t)stream;
    const int R = (int)rows, Cc = (int)cols, lg = take_log ? 1 : 0;
    const size_t elem = raw_dtype == 0 ? 2 : 4;
    if (transpose && cols % 8 == 0 && rows % 4 == 0 && ((uintptr_t)raw % 16) == 0 && ((uintptr_t)out % 16) == 0) {
        constexpr int NV = 8;
        const dim3 vgrid((unsigned)tiles, (unsigned)ceil_div(A, (int64_t)NV));
        if (raw_dtype == 0)
            preprocess_views_kernel<uint16_t, NV><<<vgrid, 256, 0, s>>>((const uint16_t*)raw, dark, flat, (int)A, R, Cc,
                                                                       min_transmission, lg, out);
        else
            preprocess_views_kernel<float, NV><<<vgrid, 256, 0, s>>>((const float*)raw, dark, flat, (int)A, R, Cc,
                                                                    min_transmission, lg, out);
        CBCT_LAUNCHED("preprocess kernel launch");
        return CBCT_OK;
    }
    if (transpose && rows % 2 == 0 && cols % 2 == 0 && ((uintptr_t)raw % (2 * elem)) == 0 &&
        ((uintptr_t)out % 8) == 0) {
        const dim3 pgrid((unsigned)(ceil_div(rows, (int64_t)PT) * ceil_div(cols, (int64_t)PW)), (unsigned)A);
        if (raw_dtype == 0)
            preprocess_pair_kernel<uint16_t><<<pgrid, 256, 0, s>>>((const uint16_t*)raw, dark, flat, R, Cc,
                                                                  min_transmission, lg, out);
        else
            preprocess_pair_kernel<float><<<pgrid, 256, 0, s>>>((const float*)raw, dark, flat, R, Cc,
                                                               min_transmission, lg, out);
        CBCT_LAUNCHED("preprocess kernel launch");
        return CBCT_OK;
    }
    if (raw_dtype == 0) {
        const uint16_t* r = (const uint16_t*)raw;
        if (transpose) preprocess_kernel<uint16_t, 1><<<grid, 256, 0, s>>>(r, dark, flat, A, R, Cc, min_transmission, lg, out);
        else preprocess_kernel<uint16_t, 0><<<grid, 256, 0, s>>>(r, dark, flat, A, R, Cc, min_transmission, lg, out);
    } else {
        const float* r = (const float*)raw;
        if (transpose) preprocess_kernel<float, 1><<<grid, 256, 0, s>>>(r, dark, flat, A, R, Cc, min_transmission, lg, out);
        else preprocess_kernel<float, 0><<<grid, 256, 0, s>>>(r, dark, flat, A, R, Cc, min_transmission, lg, out);
    }
    CBCT_LAUNCHED("preprocess kernel launch");
    return CBCT_OK;
}

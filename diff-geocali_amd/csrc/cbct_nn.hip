// BF16 MFMA kernels of the learned geometry warm start (gfx950).
//
// Every layer of the network is a GEMM on the matrix cores (v_mfma_f32_16x16x32_bf16:
// bf16 operands, FP32 accumulation) or a small memory-bound gather:
//   3x3 convolution        im2col (bias column included) -> GEMM
//   its data gradient      GEMM -> col2im (a gather, no atomics)
//   its weight gradient    GEMM with the rows as the (split) reduction dimension
//   linear layers          GEMM with a bias/ReLU epilogue; bias gradient = column sum
// Reductions that produce weights, losses or gradients run in a fixed order, so a
// training run is bitwise reproducible.
#include "cbct_common.h"
#include "cbct_nn.h"

#include <algorithm>

namespace cbct {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16, round to nearest even (NaN stays NaN, infinities stay infinite)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

constexpr int GB_M = 64, GB_N = 64, GB_K = 32;
constexpr int LDK = GB_K + 8;        // LDS row stride in bf16 elements (80 bytes, 16-byte aligned rows)

struct GemmArgs {
    const uint16_t* A;
    const uint16_t* B;
    const float* bias;
    float* C;
    uint16_t* Cb;
    float* partial;
    int64_t M, N, K, lda, ldb, ldc, ldcb, kchunk;
    int act, splits;
};

__device__ __forceinline__ void gemm_store(const GemmArgs& g, int64_t m, int64_t n, float val) {
    if (g.bias) val += g.bias[n];
    if (g.act) val = fmaxf(val, 0.0f);
    if (g.C) g.C[m * g.ldc + n] = val;
    if (g.Cb) g.Cb[m * g.ldcb + n] = f2bf(val);
}

// One 64x64 output tile per workgroup of 4 waves (each wave a 32x32 quarter = 2x2 MFMA
// tiles of 16x16); K in steps of 32 staged through LDS as [row][k] for A and [col][k] for B,
// so every MFMA operand is one 16-byte LDS read.  blockIdx.z selects the K range.
template <int TA, int TB>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t As[GB_M * LDK];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[GB_N * LDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t m0 = (int64_t)blockIdx.x * GB_M, n0 = (int64_t)blockIdx.y * GB_N;
    const int64_t kbeg = (int64_t)blockIdx.z * g.kchunk;
    const int64_t kend = min(g.K, kbeg + g.kchunk);
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int64_t k0 = kbeg; k0 < kend; k0 += GB_K) {
        {   // A tile (64 rows x 32 k): one 16-byte global load per thread
            uint4 v = {0u, 0u, 0u, 0u};
            if (!TA) {
                const int r = tid >> 2, kc = (tid & 3) * 8;
                const int64_t m = m0 + r, k = k0 + kc;
                if (m < g.M && k < kend) v = *reinterpret_cast<const uint4*>(g.A + m * g.lda + k);
                *reinterpret_cast<uint4*>(As + r * LDK + kc) = v;
            } else {
                const int kr = tid >> 3, mc = (tid & 7) * 8;
                const int64_t k = k0 + kr, m = m0 + mc;
                if (k < kend && m < g.M) v = *reinterpret_cast<const uint4*>(g.A + k * g.lda + m);
                const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
                for (int i = 0; i < 8; ++i) As[(mc + i) * LDK + kr] = e[i];
            }
        }
        {   // B tile stored as [n][k]
            uint4 v = {0u, 0u, 0u, 0u};
            if (TB) {
                const int r = tid >> 2, kc = (tid & 3) * 8;
                const int64_t n = n0 + r, k = k0 + kc;
                if (n < g.N && k < kend) v = *reinterpret_cast<const uint4*>(g.B + n * g.ldb + k);
                *reinterpret_cast<uint4*>(Bs + r * LDK + kc) = v;
            } else {
                const int kr = tid >> 3, nc = (tid & 7) * 8;
                const int64_t k = k0 + kr, n = n0 + nc;
                if (k < kend && n < g.N) v = *reinterpret_cast<const uint4*>(g.B + k * g.ldb + n);
                const uint16_t* e = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
                for (int i = 0; i < 8; ++i) Bs[(nc + i) * LDK + kr] = e[i];
            }
        }
        __syncthreads();
        // operand lanes: row (or column) = lane % 16, eight consecutive k from 8 * (lane / 16)
        const int rr = lane & 15, koff = (lane >> 4) * 8;
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8*>(As + (wm + i * 16 + rr) * LDK + koff);
#pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf16x8*>(Bs + (wn + j * 16 + rr) * LDK + koff);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // accumulator lanes: column = lane % 16, rows 4 * (lane / 16) + r
    const int cc = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t m = m0 + wm + i * 16 + rq + r, n = n0 + wn + j * 16 + cc;
                if (m < g.M && n < g.N) {
                    if (g.splits > 1) g.partial[((int64_t)blockIdx.z * g.M + m) * g.N + n] = acc[i][j][r];
                    else gemm_store(g, m, n, acc[i][j][r]);
                }
            }
}

__global__ void gemm_reduce_kernel(GemmArgs g) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t MN = g.M * g.N;
    if (i >= MN) return;
    float s = 0.f;
    for (int z = 0; z < g.splits; ++z) s += g.partial[(int64_t)z * MN + i];
    gemm_store(g, i / g.N, i % g.N, s);
}

__global__ void im2col_kernel(const uint16_t* __restrict__ x, int64_t N, int H, int W, int C, int stride,
                              int Ho, int Wo, int kpad, uint16_t* __restrict__ col) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int kch = kpad >> 3;
    if (idx >= N * Ho * Wo * kch) return;
    const int kc = (int)(idx % kch);
    const int64_t row = idx / kch;
    const int ox = (int)(row % Wo);
    const int64_t t = row / Wo;
    const int oy = (int)(t % Ho);
    const int64_t n = t / Ho;
    const int k0 = kc * 8, KC = 9 * C;
    uint16_t* dst = col + row * kpad + k0;
    if ((C & 7) == 0 && k0 + 8 <= KC) {     // eight channels of one tap: one 16-byte copy
        const int tap = k0 / C, c = k0 - tap * C;
        const int iy = oy * stride + tap / 3 - 1, ix = ox * stride + tap % 3 - 1;
        uint4 v = {0u, 0u, 0u, 0u};
        if (iy >= 0 && iy < H && ix >= 0 && ix < W)
            v = *reinterpret_cast<const uint4*>(x + ((n * H + iy) * W + ix) * C + c);
        *reinterpret_cast<uint4*>(dst) = v;
        return;
    }
    uint16_t e[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int k = k0 + i;
        uint16_t val = 0;
        if (k < KC) {
            const int tap = k / C, c = k - tap * C;
            const int iy = oy * stride + tap / 3 - 1, ix = ox * stride + tap % 3 - 1;
            if (iy >= 0 && iy < H && ix >= 0 && ix < W) val = x[((n * H + iy) * W + ix) * C + c];
        } else if (k == KC) {
            val = 0x3f80u;                  // 1.0: the bias column
        }
        e[i] = val;
    }
    *reinterpret_cast<uint4*>(dst) = *reinterpret_cast<const uint4*>(e);
}

__global__ void col2im_kernel(const float* __restrict__ dcol, int64_t N, int H, int W, int C, int stride,
                              int Ho, int Wo, int kpad, float* __restrict__ dx, uint16_t* __restrict__ dxb) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N * H * W * C) return;
    const int c = (int)(idx % C);
    int64_t t = idx / C;
    const int ix = (int)(t % W);
    t /= W;
    const int iy = (int)(t % H);
    const int64_t n = t / H;
    float s = 0.f;
    for (int ky = 0; ky < 3; ++ky) {
        const int ty = iy + 1 - ky;
        if (ty < 0 || ty % stride) continue;
        const int oy = ty / stride;
        if (oy >= Ho) continue;
        for (int kx = 0; kx < 3; ++kx) {
            const int tx = ix + 1 - kx;
            if (tx < 0 || tx % stride) continue;
            const int ox = tx / stride;
            if (ox >= Wo) continue;
            s += dcol[((n * Ho + oy) * Wo + ox) * kpad + (ky * 3 + kx) * C + c];
        }
    }
    if (dx) dx[idx] = s;
    if (dxb) dxb[idx] = f2bf(s);
}

__global__ void relu_backward_kernel(const float* __restrict__ dy, const float* __restrict__ y, int64_t n,
                                     float* __restrict__ dx, uint16_t* __restrict__ dxb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = y[i] > 0.f ? dy[i] : 0.f;
    if (dx) dx[i] = v;
    if (dxb) dxb[i] = f2bf(v);
}

__global__ void spatial_mean_kernel(const float* __restrict__ x, int64_t R, int S, int C, float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R * C) return;
    const int64_t r = i / C;
    const int c = (int)(i % C);
    float s = 0.f;
    for (int k = 0; k < S; ++k) s += x[(r * S + k) * C + c];
    out[i] = s / (float)S;
}

__global__ void spatial_mean_backward_kernel(const float* __restrict__ dout, int64_t R, int S, int C,
                                             float* __restrict__ dx, uint16_t* __restrict__ dxb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R * S * C) return;
    const int c = (int)(i % C);
    const int64_t r = i / ((int64_t)S * C);
    const float v = dout[r * C + c] / (float)S;
    if (dx) dx[i] = v;
    if (dxb) dxb[i] = f2bf(v);
}

__global__ void view_concat_kernel(const float* __restrict__ feat, int64_t P, int A, int C, uint16_t* __restrict__ h) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P * C) return;
    const int64_t p = i / C;
    const int c = (int)(i % C);
    const float* f = feat + p * A * C + c;
    float s = 0.f;
    for (int a = 0; a < A; ++a) s += f[(int64_t)a * C];
    const uint16_t mean = f2bf(s / (float)A);
    for (int a = 0; a < A; ++a) {
        uint16_t* row = h + (p * A + a) * 2 * C;
        row[c] = f2bf(f[(int64_t)a * C]);
        row[C + c] = mean;
    }
}

__global__ void view_concat_backward_kernel(const float* __restrict__ dh, int64_t P, int A, int C,
                                            float* __restrict__ dfeat) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P * C) return;
    const int64_t p = i / C;
    const int c = (int)(i % C);
    float s = 0.f;
    for (int a = 0; a < A; ++a) s += dh[(p * A + a) * 2 * C + C + c];
    const float share = s / (float)A;
    for (int a = 0; a < A; ++a) dfeat[(p * A + a) * C + c] = dh[(p * A + a) * 2 * C + c] + share;
}

__global__ void colsum_kernel(const float* __restrict__ x, int64_t R, int64_t C, float* __restrict__ out) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    float s = 0.f;
    for (int64_t r = 0; r < R; ++r) s += x[r * C + c];
    out[c] = s;
}

__global__ __launch_bounds__(256) void mse_kernel(const float* __restrict__ pred, const float* __restrict__ target,
                                                  int64_t n, float* __restrict__ loss, float* __restrict__ dpred) {
    float s = 0.f;
    const float inv = 1.0f / (float)n;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const float d = pred[i] - target[i];
        s = fmaf(d, d, s);
        dpred[i] = 2.0f * d * inv;
    }
    __shared__ float red[4];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) loss[0] = ((red[0] + red[1]) + (red[2] + red[3])) * inv;
}

__global__ void adam_kernel(float* __restrict__ w, uint16_t* __restrict__ wb, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, int64_t n, float lr, float b1, float b2,
                            float eps, float bc1, float bc2, float wd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps) - lr * wd * w[i];
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

__global__ void to_bf16_kernel(const float* __restrict__ x, int64_t n, uint16_t* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

__global__ void prepare_input_kernel(const float* __restrict__ meas, const float* __restrict__ nom, int64_t A,
                                     int U, int V, int f, float scale, float gain, uint16_t* __restrict__ out) {
    const int Uo = U / f, Vo = V / f;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A * Uo * Vo) return;
    const int vo = (int)(i % Vo);
    const int64_t t = i / Vo;
    const int uo = (int)(t % Uo);
    const int64_t a = t / Uo;
    float sm = 0.f, sn = 0.f;
    for (int du = 0; du < f; ++du) {
        const int64_t base = (a * U + uo * f + du) * V + (int64_t)vo * f;
        for (int dv = 0; dv < f; ++dv) {
            sm += meas[base + dv];
            sn += nom[base + dv];
        }
    }
    const float w = scale / (float)(f * f);
    out[2 * i] = f2bf(sn * w);
    out[2 * i + 1] = f2bf((sm - sn) * (w * gain));
}

// One workgroup: thread t takes rows t, t+256, ...; per row e = pred - target, d = J e,
// dpred = 2 J^T d / (R Q); the squared norms are summed per thread and then over the
// workgroup in a fixed order.
__global__ __launch_bounds__(256) void jacobian_mse_kernel(const float* __restrict__ pred,
                                                           const float* __restrict__ target,
                                                           const float* __restrict__ J, int64_t jstride, int64_t R,
                                                           int K, int Q, float* __restrict__ loss,
                                                           float* __restrict__ dpred) {
    const float inv = 2.0f / ((float)R * (float)Q);
    float s = 0.f;
    for (int64_t r = threadIdx.x; r < R; r += 256) {
        float e[16], g[16];
        for (int k = 0; k < K; ++k) {
            e[k] = pred[r * K + k] - target[r * K + k];
            g[k] = 0.f;
        }
        const float* Jr = J + r * jstride;
        for (int q = 0; q < Q; ++q) {
            float d = 0.f;
            for (int k = 0; k < K; ++k) d = fmaf(Jr[q * K + k], e[k], d);
            s = fmaf(d, d, s);
            for (int k = 0; k < K; ++k) g[k] = fmaf(Jr[q * K + k], d, g[k]);
        }
        for (int k = 0; k < K; ++k) dpred[r * K + k] = g[k] * inv;
    }
    __shared__ float red[4];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) loss[0] = ((red[0] + red[1]) + (red[2] + red[3])) * (0.5f * inv);
}

constexpr int SQ_BLOCKS = 512;

// Stage 1: block b sums (a-b)^2 over the indices i = b*256 + t + k*SQ_BLOCKS*256 (fixed order).
__global__ __launch_bounds__(256) void sqdiff_partial_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                                             int64_t n, float* __restrict__ part) {
    float s = 0.f;
    const int64_t stride = (int64_t)SQ_BLOCKS * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
        const float d = a[i] - b[i];
        s = fmaf(d, d, s);
    }
    __shared__ float red[4];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void sqdiff_final_kernel(const float* __restrict__ part, int64_t n,
                                                           float* __restrict__ out) {
    float s = 0.f;
    for (int i = threadIdx.x; i < SQ_BLOCKS; i += 256) s += part[i];
    __shared__ float red[4];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = ((red[0] + red[1]) + (red[2] + red[3])) / (float)n;
}

unsigned blocks_for(int64_t n, int threads = 256) { return (unsigned)ceil_div(n, (int64_t)threads); }

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace
}  // namespace cbct

using namespace cbct;

extern "C" size_t cbct_nn_gemm_workspace_bytes(int64_t M, int64_t N, int splits) {
    if (splits <= 1 || M <= 0 || N <= 0) return 0;
    return (size_t)splits * (size_t)M * (size_t)N * sizeof(float);
}

extern "C" int cbct_nn_gemm(int trans_a, int trans_b, int64_t M, int64_t N, int64_t K, const uint16_t* A,
                            int64_t lda, const uint16_t* B, int64_t ldb, const float* bias, int act, float* C,
                            int64_t ldc, uint16_t* Cb, int64_t ldcb, int splits, void* workspace,
                            size_t workspace_bytes, void* stream) {
    CBCT_CHECK_ARG(M > 0 && N > 0 && K > 0, "GEMM sizes must be positive");
    CBCT_CHECK_ARG(K % 8 == 0 || (trans_a && !trans_b),
                   "GEMM K must be a multiple of 8 unless both operands are stored K-major");
    CBCT_CHECK_ARG(!trans_a || M % 8 == 0, "GEMM M must be a multiple of 8 when A is transposed");
    CBCT_CHECK_ARG(trans_b || N % 8 == 0, "GEMM N must be a multiple of 8 when B is not transposed");
    CBCT_CHECK_ARG(lda % 8 == 0 && ldb % 8 == 0, "GEMM leading dimensions must be multiples of 8");
    CBCT_CHECK_ARG(lda >= (trans_a ? M : K) && ldb >= (trans_b ? K : N), "GEMM leading dimension too small");
    CBCT_CHECK_ARG(A && B && aligned16(A) && aligned16(B), "GEMM operands must be 16-byte aligned");
    CBCT_CHECK_ARG(C || Cb, "GEMM needs an output");
    CBCT_CHECK_ARG((!C || ldc >= N) && (!Cb || ldcb >= N), "GEMM output leading dimension too small");
    CBCT_CHECK_ARG(act == 0 || act == 1, "act must be 0 (identity) or 1 (ReLU)");
    CBCT_CHECK_ARG(splits >= 1 && splits <= 4096, "splits must be in [1, 4096]");
    CBCT_CHECK_ARG(ceil_div(N, (int64_t)GB_N) <= 65535 && ceil_div(M, (int64_t)GB_M) <= INT32_MAX,
                   "GEMM sizes too large for one launch");
    GemmArgs g;
    g.A = A; g.B = B; g.bias = bias; g.C = C; g.Cb = Cb; g.partial = nullptr;
    g.M = M; g.N = N; g.K = K; g.lda = lda; g.ldb = ldb; g.ldc = ldc; g.ldcb = ldcb;
    g.act = act;
    // K range of each split: a multiple of the K step
    const int64_t steps = ceil_div(K, (int64_t)GB_K);
    const int64_t per = ceil_div(steps, (int64_t)splits);
    g.kchunk = per * GB_K;
    g.splits = (int)ceil_div(K, g.kchunk);
    if (g.splits > 1) {
        CBCT_CHECK_ARG(workspace && workspace_bytes >= cbct_nn_gemm_workspace_bytes(M, N, g.splits),
                       "workspace too small for cbct_nn_gemm");
        g.partial = (float*)workspace;
    }
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid((unsigned)ceil_div(M, (int64_t)GB_M), (unsigned)ceil_div(N, (int64_t)GB_N), (unsigned)g.splits);
    if (!trans_a && !trans_b) gemm_kernel<0, 0><<<grid, 256, 0, s>>>(g);
    else if (!trans_a && trans_b) gemm_kernel<0, 1><<<grid, 256, 0, s>>>(g);
    else if (trans_a && !trans_b) gemm_kernel<1, 0><<<grid, 256, 0, s>>>(g);
    else gemm_kernel<1, 1><<<grid, 256, 0, s>>>(g);
    CBCT_LAUNCHED("GEMM kernel launch");
    if (g.splits > 1) {
        gemm_reduce_kernel<<<blocks_for(M * N), 256, 0, s>>>(g);
        CBCT_LAUNCHED("GEMM split-K reduce launch");
    }
    return CBCT_OK;
}

static int check_conv(int64_t N, int64_t H, int64_t W, int64_t C, int stride, int64_t kpad) {
    CBCT_CHECK_ARG(N > 0 && H > 0 && W > 0 && C > 0, "convolution sizes must be positive");
    CBCT_CHECK_ARG(stride == 1 || stride == 2, "stride must be 1 or 2");
    CBCT_CHECK_ARG(kpad >= 9 * C + 1 && kpad % 8 == 0, "kpad must be a multiple of 8 and at least 9*C+1");
    CBCT_CHECK_ARG(H < (1 << 20) && W < (1 << 20) && kpad < (1 << 24), "convolution sizes too large");
    return CBCT_OK;
}

extern "C" int cbct_nn_im2col(const uint16_t* x, int64_t N, int64_t H, int64_t W, int64_t C, int stride,
                              int64_t kpad, uint16_t* col, void* stream) {
    int st;
    if ((st = check_conv(N, H, W, C, stride, kpad))) return st;
    CBCT_CHECK_ARG(x && col && aligned16(x) && aligned16(col), "im2col buffers must be 16-byte aligned");
    const int Ho = (int)((H - 1) / stride + 1), Wo = (int)((W - 1) / stride + 1);
    const int64_t n = N * Ho * Wo * (kpad / 8);
    im2col_kernel<<<blocks_for(n), 256, 0, (hipStream_t)stream>>>(x, N, (int)H, (int)W, (int)C, stride, Ho, Wo,
                                                                   (int)kpad, col);
    CBCT_LAUNCHED("im2col kernel launch");
    return CBCT_OK;
}

extern "C" int cbct_nn_col2im(const float* dcol, int64_t N, int64_t H, int64_t W, int64_t C, int stride,
                              int64_t kpad, float* dx, uint16_t* dxb, void* stream) {
    int st;
    if ((st = check_conv(N, H, W, C, stride, kpad))) return st;
    CBCT_CHECK_ARG(dcol && (dx || dxb), "col2im needs an input and an output");
    const int Ho = (int)((H - 1) / stride + 1), Wo = (int)((W - 1) / stride + 1);
    col2im_kernel<<<blocks_for(N * H * W * C), 256, 0, (hipStream_t)stream>>>(dcol, N, (int)H, (int)W, (int)C,
                                                                               stride, Ho, Wo, (int)kpad, dx, dxb);
    CBCT_LAUNCHED("col2im kernel launch");
    return CBCT_OK;
}

extern "C" int cbct_nn_relu_backward(const float* dy, const float* y, int64_t n, float* dx, uint16_t* dxb,
                                     void* stream) {
    CBCT_CHECK_ARG(n > 0 && dy && y && (dx || dxb), "relu_backward needs n > 0, inputs and an output");
    relu_backward_kernel<<<blocks_for(n), 256, 0, (hipStream_t)stream>>>(dy, y, n, dx, dxb);
    CBCT_LAUNCHED("ReLU backward kernel launch");
    return CBCT_OK;
}

extern "C" int cbct_nn_spatial_mean(const float* x, int64_t R, int64_t S, int64_t C, float* out, void* stream) {
    CBCT_CHECK_ARG(R > 0 && S > 0 && C > 0 && x && out, "spatial_mean needs positive sizes and buffers");
    spatial_mean_kernel<<<blocks_for(R * C), 256, 0, (hipStream_t)stream>>>(x, R, (int)S, (int)C, out);
    CBCT_LAUNCHED("spatial mean kernel launch");
    return CBCT_OK;
}

extern "C" int cbct_nn_spatial_mean_backward(const float* dout, int64_t R, int64_t S, int64_t C, float* dx,
                                             uint16_t* dxb, void* stream) {
    CBCT_CHECK_ARG(R > 0 && S > 0 && C > 0 && dout && (dx || dxb), "spatial_mean_backward needs sizes and buffers");
    spatial_mean_backward_kernel<<<blocks_for(R * S * C), 256, 0, (hipStream_t)stream>>>(dout, R, (int)S, (int)C,
                                                                                         dx, dxb);
    CBCT_LAUNCHED("spatial mean backward kernel launch");
    return CBCT_OK;
}

extern "C" int cbct_nn_view_concat(const float* feat, int64_t P, int64_t A, int64_t C, uint16_t* h, void* stream) {
    CBCT_CHECK_ARG(P > 0 && A > 0 && C > 0 && feat && h, "view_concat needs positive sizes and buffers");
    view_concat_kernel<<<blocks_for(P * C, 64), 64, 0, (hipStream_t)stream>>>(feat, P, (int)A, (int)C, h);
    CBCT_LAUNCHED("view concat kernel launch");
    return CBCT_OK;
}

extern "C" int cbct_nn_view_concat_backward(const float* dh, int64_t P, int64_t A, int64_t C, float* dfeat,
                                            void* stream) {
    CBCT_CHECK_ARG(P > 0 && A > 0 && C > 0 && dh && dfeat, "view_concat_backward needs positive sizes and buffers");
    view_concat_backward_kernel<<<blocks_for(P * C, 64), 64, 0, (hipStream_t)stream>>>(dh, P, (int)A, (int)C, dfeat);
    CBCT_LAUNCHED("view concat backward kernel launch");
    return CBCT_OK;
}

extern "C" int cbct_nn_colsum(const float* x, int64_t R, int64_t C, float* out, void* stream) {
    CBCT_CHECK_ARG(R > 0 && C > 0 && x && out, "colsum needs positive sizes and buffers");
    colsum_kernel<<<blocks_for(C, 64), 64, 0, (hipStream_t)stream>>>(x, R, C, out);
    CBCT_LAUNCHED("column sum kernel launch");
    return CBCT_OK;
}

extern "C" int cbct_nn_mse(const float* pred, const float* target, int64_t n, float* loss, float* dpred,
                           void* stream) {
    CBCT_CHECK_ARG(n > 0 && pred && target && loss && dpred, "mse needs n > 0 and buffers");
    mse_kernel<<<1, 256, 0, (hipStream_t)stream>>>(pred, target, n, loss, dpred);
    CBCT_LAUNCHED("MSE kernel launch");
    return CBCT_OK;
}

extern "C" int cbct_nn_adam(float* w, uint16_t* wb, const float* g, float* m, float* v, int64_t n, float lr,
                            float beta1, float beta2, float eps, float bc1, float bc2, float weight_decay,
                            void* stream) {
    CBCT_CHECK_ARG(n > 0 && w && g && m && v, "adam needs n > 0 and buffers");
    CBCT_CHECK_ARG(bc1 > 0.f && bc2 > 0.f && eps > 0.f, "adam bias corrections and eps must be positive");
    adam_kernel<<<blocks_for(n), 256, 0, (hipStream_t)stream>>>(w, wb, g, m, v, n, lr, beta1, beta2, eps, bc1, bc2,
                                                                weight_decay);
    CBCT_LAUNCHED("Adam kernel launch");
    return CBCT_OK;
}

extern "C" int cbct_nn_to_bf16(const float* x, int64_t n, uint16_t* y, void* stream) {
    CBCT_CHECK_ARG(n > 0 && x && y, "to_bf16 needs n > 0 and buffers");
    to_bf16_kernel<<<blocks_for(n), 256, 0, (hipStream_t)stream>>>(x, n, y);
    CBCT_LAUNCHED("bf16 conversion kernel launch");
    return CBCT_OK;
}

extern "C" int cbct_nn_prepare_input(const float* measured, const float* nominal, int64_t A, int64_t U, int64_t V,
                                     int factor, float scale, float diff_gain, uint16_t* out, void* stream) {
    CBCT_CHECK_ARG(A > 0 && U > 0 && V > 0 && factor >= 1, "prepare_input needs positive sizes");
    CBCT_CHECK_ARG(U % factor == 0 && V % factor == 0, "U and V must be multiples of the downsampling factor");
    CBCT_CHECK_ARG(measured && nominal && out, "prepare_input needs buffers");
    const int64_t n = A * (U / factor) * (V / factor);
    prepare_input_kernel<<<blocks_for(n), 256, 0, (hipStream_t)stream>>>(measured, nominal, A, (int)U, (int)V, factor,
                                                                         scale, diff_gain, out);
    CBCT_LAUNCHED("input preparation kernel launch");
    return CBCT_OK;
}

extern "C" int cbct_nn_jacobian_mse(const float* pred, const float* target, const float* J, int64_t jstride,
                                    int64_t R, int64_t K, int64_t Q, float* loss, float* dpred, void* stream) {
    CBCT_CHECK_ARG(R > 0 && K > 0 && K <= 16 && Q > 0, "jacobian_mse needs R > 0, 0 < K <= 16 and Q > 0");
    CBCT_CHECK_ARG(jstride == 0 || jstride >= Q * K, "jstride must be 0 (shared J) or at least Q*K");
    CBCT_CHECK_ARG(pred && target && J && loss && dpred, "jacobian_mse needs buffers");
    jacobian_mse_kernel<<<1, 256, 0, (hipStream_t)stream>>>(pred, target, J, jstride, R, (int)K, (int)Q, loss, dpred);
    CBCT_LAUNCHED("Jacobian MSE kernel launch");
    return CBCT_OK;
}

extern "C" size_t cbct_nn_sqdiff_workspace_bytes(void) { return SQ_BLOCKS * sizeof(float); }

extern "C" int cbct_nn_sqdiff_mean(const float* a, const float* b, int64_t n, float* out, void* workspace,
                                   size_t workspace_bytes, void* stream) {
    CBCT_CHECK_ARG(n > 0 && a && b && out, "sqdiff_mean needs n > 0 and buffers");
    CBCT_CHECK_ARG(workspace && workspace_bytes >= cbct_nn_sqdiff_workspace_bytes(),
                   "workspace too small for cbct_nn_sqdiff_mean");
    float* part = (float*)workspace;
    sqdiff_partial_kernel<<<SQ_BLOCKS, 256, 0, (hipStream_t)stream>>>(a, b, n, part);
    CBCT_LAUNCHED("squared-difference kernel launch");
    sqdiff_final_kernel<<<1, 256, 0, (hipStream_t)stream>>>(part, n, out);
    CBCT_LAUNCHED("squared-difference final kernel launch");
    return CBCT_OK;
}

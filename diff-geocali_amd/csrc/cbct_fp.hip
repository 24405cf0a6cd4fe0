// Forward ray-integral projector and its adjoint (volume gradient plus the
// eight per-view geometry gradients) for gfx950.
//
// Ray assignment: a workgroup of 256 lanes owns a 16x16 tile of detector
// pixels of one view; each wave owns an 8x8 sub-tile (8 consecutive v per
// row, 8 rows of u).  Neighbouring lanes march neighbouring rays in lockstep,
// so one wave-wide gather touches a compact 3-D neighbourhood of the volume
// and mostly hits the vector L1 / L2 instead of HBM.
//
// Sample schedule (identical to the reference, csrc/src_cbct_forwardprojector_cuda.cu:321):
// t starts at the box entry and is advanced by float addition of step_size
// while t <= t_max.
#include "cbct_common.h"
#include "cbct_fp_slab.h"

#include <algorithm>
#include <mutex>
#include <string>

namespace cbct {

constexpr int FP_TILE = 16;          // rays per tile edge
constexpr int FP_THREADS = FP_TILE * FP_TILE;

struct TileMap {
    int a, u, v;
};

// Block -> (view, tile) and lane -> ray.  Blocks are dispatched round-robin
// over the 8 XCDs; consecutive tiles of one view are kept on one XCD so that
// the rays sharing an L2 are neighbours.
__device__ __forceinline__ TileMap tile_map(int tiles_u, int tiles_v, int nblocks) {
    const int xcd = blockIdx.x & 7;
    const int per_xcd = (nblocks + 7) >> 3;
    int tile = xcd * per_xcd + (blockIdx.x >> 3);
    if ((nblocks & 7) != 0) tile = blockIdx.x;   // uneven grid: plain order
    const int tv = tile % tiles_v;
    const int rest = tile / tiles_v;
    const int tu = rest % tiles_u;
    TileMap m;
    m.a = rest / tiles_u;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    m.v = tv * FP_TILE + ((wave & 1) << 3) + (lane & 7);
    m.u = tu * FP_TILE + ((wave >> 1) << 3) + (lane >> 3);
    return m;
}

template <bool DIV>
__global__ __launch_bounds__(FP_THREADS) void fp_forward_kernel(
    const float* __restrict__ vol, Box b, const ViewFrame* __restrict__ frames,
    int U, int V, int tiles_u, int tiles_v, int nblocks,
    float pix, float step, float vs, float inv_vs, float* __restrict__ proj) {
    const TileMap m = tile_map(tiles_u, tiles_v, nblocks);
    if (m.u >= U || m.v >= V) return;
    const ViewFrame f = frames[m.a];
    float integral = 0.0f;
    Ray ray;
    if (make_ray(f, m.u, m.v, pix, ray)) {
        float tmin, tmax;
        int entry;
        if (intersect(b, f, ray, tmin, tmax, entry)) {
            for (float t = tmin; t <= tmax; t += step) {
                const float px = fmaf(t, ray.dx, f.ox);
                const float py = fmaf(t, ray.dy, f.oy);
                const float pz = fmaf(t, ray.dz, f.oz);
                const float x = to_voxel<DIV>(px, b.offx, vs, inv_vs, b.hx);
                const float y = to_voxel<DIV>(py, b.offy, vs, inv_vs, b.hy);
                const float z = to_voxel<DIV>(pz, b.offz, vs, inv_vs, b.hz);
                const Corners c = corners(b, x, y, z);
                const Taps tp = load_taps(vol, c);
                integral += tri_value(tp, c.fx, c.fy, c.fz);
            }
            integral *= step;
        }
    }
    proj[((int64_t)m.a * U + m.u) * V + m.v] = integral;
}

// Per-ray adjoint.  Writes one [8] partial of the geometry gradients per
// workgroup (its tile); fp_reduce_geometry sums them in a fixed order.
template <bool GVOL>
__global__ __launch_bounds__(FP_THREADS) void fp_backward_kernel(
    const float* __restrict__ grad, const float* __restrict__ vol, Box b,
    const ViewFrame* __restrict__ frames, int U, int V, int tiles_u, int tiles_v, int nblocks,
    float pix, float step, float inv_vs, float* __restrict__ gvol, float* __restrict__ partials) {
    const TileMap m = tile_map(tiles_u, tiles_v, nblocks);
    const bool active = m.u < U && m.v < V;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const float go = active ? grad[((int64_t)m.a * U + m.u) * V + m.v] : 0.0f;
    const ViewFrame f = frames[m.a];
    Ray ray;
    if (active && go != 0.0f && make_ray(f, m.u, m.v, pix, ray)) {
        float gox = 0.f, goy = 0.f, goz = 0.f, gdx = 0.f, gdy = 0.f, gdz = 0.f;
        float tmin, tmax;
        int entry;
        if (intersect(b, f, ray, tmin, tmax, entry)) {
            const float gs = go * step;
            float gtmin = 0.0f;
            for (float t = tmin; t <= tmax; t += step) {
                const float px = fmaf(t, ray.dx, f.ox);
                const float py = fmaf(t, ray.dy, f.oy);
                const float pz = fmaf(t, ray.dz, f.oz);
                const float x = fmaf(px - b.offx, inv_vs, b.hx);
                const float y = fmaf(py - b.offy, inv_vs, b.hy);
                const float z = fmaf(pz - b.offz, inv_vs, b.hz);
                const Corners c = corners(b, x, y, z);
                const Taps tp = load_taps(vol, c);
                if (GVOL) {
                    const Weights w = tri_weights(c.fx, c.fy, c.fz);
                    float* p = gvol + c.i000;
                    atomicAdd(p, gs * w.w000);
                    atomicAdd(p + c.dx1, gs * w.w100);
                    atomicAdd(p + c.dy1, gs * w.w010);
                    atomicAdd(p + c.dy1 + c.dx1, gs * w.w110);
                    atomicAdd(p + c.dz1, gs * w.w001);
                    atomicAdd(p + c.dz1 + c.dx1, gs * w.w101);
                    atomicAdd(p + c.dz1 + c.dy1, gs * w.w011);
                    atomicAdd(p + c.dz1 + c.dy1 + c.dx1, gs * w.w111);
                }
                float sval, sgx, sgy, sgz;
                tri_eval(tp, c.fx, c.fy, c.fz, sval, sgx, sgy, sgz);
                // unscaled sums; the constant factor gs / voxel size is applied once below
                gox += sgx; goy += sgy; goz += sgz;
                gdx = fmaf(sgx, t, gdx); gdy = fmaf(sgy, t, gdy); gdz = fmaf(sgz, t, gdz);
            }
            const float sc = gs * inv_vs;
            gtmin = sc * fmaf(ray.dz, goz, fmaf(ray.dy, goy, ray.dx * gox));
            gox *= sc; goy *= sc; goz *= sc; gdx *= sc; gdy *= sc; gdz *= sc;
            // The samples are anchored at the entry face (:454-467).
            if (entry == 0) { gox -= gtmin / ray.dx; gdx -= gtmin * tmin / ray.dx; }
            else if (entry == 1) { goy -= gtmin / ray.dy; gdy -= gtmin * tmin / ray.dy; }
            else if (entry == 2) { goz -= gtmin / ray.dz; gdz -= gtmin * tmin / ray.dz; }
        }
        // d = (T - O)/|T - O|  (:470-486)
        const float dd = fmaf(ray.dz, gdz, fmaf(ray.dy, gdy, ray.dx * gdx));
        const float il = 1.0f / ray.len;
        const float grx = fmaf(-ray.dx, dd, gdx) * il;
        const float gry = fmaf(-ray.dy, dd, gdy) * il;
        const float grz = fmaf(-ray.dz, dd, gdz) * il;
        const float Ox = gox - grx, Oy = goy - gry, Oz = goz - grz;
        // per-view parameters (:495-527)
        const float dOxb = -f.sod * f.c, dOyb = -f.sod * f.s;
        const float dTxb = fmaf(-ray.up, f.s, -f.nc);
        const float dTyb = fmaf(ray.up, f.c, -f.ns);
        acc[0] = fmaf(gry, dTyb, fmaf(grx, dTxb, fmaf(Oy, dOyb, Ox * dOxb)));
        acc[1] = fmaf(grx, f.s, -(gry * f.c));
        acc[2] = fmaf(Oy + gry, f.c, (Ox + grx) * -f.s);
        acc[3] = -pix * fmaf(gry, f.s, grx * f.c);
        acc[4] = -pix * grz;
        acc[5] = Ox; acc[6] = Oy; acc[7] = Oz;
    }
    __shared__ float red[FP_THREADS / 64][8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float s = wave_sum(acc[k]);
        if (lane == 0) red[wave][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        const int k = threadIdx.x;
        const float s = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
        const int tile_in_view = (int)(((int64_t)m.u / FP_TILE) * tiles_v + m.v / FP_TILE);
        partials[((int64_t)m.a * tiles_u * tiles_v + tile_in_view) * 8 + k] = s;
    }
}

// Sum the per-tile partials of each view in a fixed order: ggeom[k][a].
__global__ __launch_bounds__(256) void fp_reduce_geometry(const float* __restrict__ partials,
                                                          int A, int tiles, float* __restrict__ ggeom) {
    const int a = blockIdx.x;
    __shared__ float sm[256][9];
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int t = threadIdx.x; t < tiles; t += 256) {
        const float* p = partials + ((int64_t)a * tiles + t) * 8;
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] += p[k];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) sm[threadIdx.x][k] = acc[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
#pragma unroll
            for (int k = 0; k < 8; ++k) sm[threadIdx.x][k] += sm[threadIdx.x + s][k];
        }
        __syncthreads();
    }
    if (threadIdx.x < 8) ggeom[(int64_t)threadIdx.x * A + a] = sm[0][threadIdx.x];
}

__global__ __launch_bounds__(FP_THREADS) void fp_count_kernel(
    Box b, const ViewFrame* __restrict__ frames, int U, int V, int tiles_u, int tiles_v, int nblocks,
    float pix, float step, unsigned long long* __restrict__ count) {
    const TileMap m = tile_map(tiles_u, tiles_v, nblocks);
    unsigned long long n = 0;
    if (m.u < U && m.v < V) {
        const ViewFrame f = frames[m.a];
        Ray ray;
        float tmin, tmax;
        int entry;
        if (make_ray(f, m.u, m.v, pix, ray) && intersect(b, f, ray, tmin, tmax, entry))
            for (float t = tmin; t <= tmax; t += step) ++n;
    }
    __shared__ unsigned long long red[FP_THREADS / 64];
    for (int s = 32; s >= 1; s >>= 1) n += __shfl_xor(n, s, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(count, (red[0] + red[1]) + (red[2] + red[3]));
}

// Process-wide tuning of the projector kernels (cbct_fp_set_tuning / cbct_fp_set_tuning_op), one
// setting for the kernels that only gather (forward, geometry-only adjoint, calibration march) and
// one for the adjoint that also scatters the volume gradient into LDS: gathers favour 8x8 lane
// blocks (neighbouring taps share banks and cache lines), the LDS integer atomics favour the
// interleaved layout (fewer lanes of a wave on one bank).  Defaults measured on the 256^3 x 720
// views, 768x640 workload: the gathering kernels' 20 KB box leaves room for 7 workgroups per CU (a
// box that fits 8 measured the same); the adjoint's two 15 KB boxes plus its reductions keep the 5
// workgroups its registers allow (two 16 KB boxes would leave room for only 4).
struct FpTuning {
    int use_slab;         // slab-staged kernels (1) or direct kernels (0)
    int T;                // slab thickness in voxel planes
    int cap;              // LDS floats per staged box
    int layout;           // lane layout of a tile
};
// The process-wide defaults (cbct_fp_set_tuning_op, cbct_fp_set_deterministic), read by the entry
// points without options and by the *_ex entry points given options == NULL.  The mutex keeps a
// concurrent set from tearing a read; a call with options never reads them.
static std::mutex g_defaults_mutex;
static FpTuning g_tuning_op[2] = {{1, 12, 5120, 0}, {1, 12, 3840, 1}};
static int g_deterministic = 0;   // volume gradient through order-independent int64 accumulation

struct FpCall {
    FpTuning t;
    int det;
};

static int check_options(const cbct_fp_options& o) {
    CBCT_CHECK_ARG(o.use_slab == 0 || o.use_slab == 1, "use_slab must be 0 or 1");
    CBCT_CHECK_ARG(o.slab_thickness >= 1 && o.slab_thickness <= 64, "slab_thickness must be in [1, 64]");
    CBCT_CHECK_ARG(o.lds_floats >= 256 && o.lds_floats <= 8192, "lds_floats must be in [256, 8192]");
    CBCT_CHECK_ARG(o.layout == 0 || o.layout == 1, "layout must be 0 or 1");
    CBCT_CHECK_ARG(o.deterministic == 0 || o.deterministic == 1, "deterministic must be 0 or 1");
    return CBCT_OK;
}

// The settings of one call in operation class op (0: gathering kernels, 1: adjoint with the
// volume gradient): the caller's options, or the process-wide defaults.
static int call_settings(const cbct_fp_options* opt, int op, FpCall& c) {
    if (opt) {
        const int st = check_options(*opt);
        if (st) return st;
        c.t = FpTuning{opt->use_slab, opt->slab_thickness, opt->lds_floats, opt->layout};
        c.det = opt->deterministic;
        return CBCT_OK;
    }
    std::lock_guard<std::mutex> lock(g_defaults_mutex);
    c.t = g_tuning_op[op];
    c.det = g_deterministic;
    return CBCT_OK;
}

int call_deterministic(const cbct_fp_options* opt, int* det) {
    FpCall c;
    const int st = call_settings(opt, 0, c);
    if (st) return st;
    *det = c.det;
    return CBCT_OK;
}

// ---- deterministic volume gradient helpers
__global__ void zero_u64_kernel(unsigned long long* __restrict__ p, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0ull;
}

// max |x| over n floats as the bit pattern of a non-negative float (order-independent max);
// one atomic per workgroup, 16-byte loads when x is 16-byte aligned.
__global__ __launch_bounds__(256) void absmax_kernel(const float* __restrict__ x, int64_t n,
                                                     unsigned* __restrict__ maxbits) {
    __shared__ float red[4];
    float m = 0.f;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t first = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n4 = ((uintptr_t)x & 15) == 0 ? n / 4 : 0;
    const float4* __restrict__ x4 = reinterpret_cast<const float4*>(x);
    for (int64_t i = first; i < n4; i += stride) {
        const float4 v = x4[i];
        m = fmaxf(m, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
    }
    for (int64_t i = 4 * n4 + first; i < n; i += stride) m = fmaxf(m, fabsf(x[i]));
    for (int s = 32; s >= 1; s >>= 1) m = fmaxf(m, __shfl_xor(m, s, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        if (m > 0.f) atomicMax(maxbits, __float_as_uint(m));
    }
}

// global exponent E: the largest per-sample contribution |g * step| < 2^40 units of 2^-E
__global__ void det_exponent_kernel(const unsigned* __restrict__ maxbits, float step, int* __restrict__ gexp) {
    const float m = __uint_as_float(maxbits[0]) * step;
    int E = 0;
    if (m > 0.f && isfinite(m)) {
        int k;
        frexpf(m, &k);                // m = f * 2^k, f in [0.5, 1)
        E = max(-60, min(120, 40 - k));
    }
    gexp[0] = E;
}

__global__ void det_convert_kernel(const unsigned long long* __restrict__ acc, int64_t n,
                                   const int* __restrict__ gexp, float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = (float)ldexp((double)(long long)acc[i], -gexp[0]);
}

int launch_zero_u64(unsigned long long* p, int64_t n, hipStream_t stream) {
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, (int64_t)256), 16384));
    zero_u64_kernel<<<blocks, 256, 0, stream>>>(p, n);
    CBCT_LAUNCHED("int64 accumulator reset");
    return CBCT_OK;
}

int launch_absmax(const float* x, int64_t n, unsigned* maxbits, hipStream_t stream) {
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, (int64_t)256 * 16), 2048));
    absmax_kernel<<<blocks, 256, 0, stream>>>(x, n, maxbits);
    CBCT_LAUNCHED("max |x| kernel");
    return CBCT_OK;
}

int launch_reduce_geometry(const float* partials, int64_t A, int64_t tiles, float* ggeom, hipStream_t stream) {
    fp_reduce_geometry<<<(unsigned)A, 256, 0, stream>>>(partials, (int)A, (int)tiles, ggeom);
    CBCT_LAUNCHED("geometry reduction kernel launch");
    return CBCT_OK;
}

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

static int64_t fp_tiles(int64_t U, int64_t V) { return ceil_div(U, FP_TILE) * ceil_div(V, FP_TILE); }

static int check_fp_common(int64_t D, int64_t H, int64_t W, int64_t A, int64_t U, int64_t V,
                           float pix, float step, float vs, float ox, float oy, float oz) {
    CBCT_CHECK_ARG(D > 0 && H > 0 && W > 0, "every volume dimension must be positive");
    CBCT_CHECK_ARG(U > 0 && V > 0, "U and V must be positive");
    CBCT_CHECK_ARG(A > 0, "angles must contain at least one view");
    // voxel offsets are 64-bit; the per-axis limits keep a plane (H*W) and a column of planes
    // (D*H) within 32 bits, and every voxel coordinate exact in FP32
    CBCT_CHECK_ARG(D < (1 << 24) && H < (1 << 24) && W < (1 << 24),
                   "every volume dimension must be below 2^24");
    CBCT_CHECK_ARG(H * W <= INT32_MAX && D * H <= INT32_MAX, "volume planes exceed the 2^31-1 element limit");
    CBCT_CHECK_ARG(A * fp_tiles(U, V) <= INT32_MAX, "too many rays for one launch");
    int st;
    if ((st = check_positive_finite(pix, "pixel_size"))) return st;
    if ((st = check_positive_finite(step, "step_size"))) return st;
    if ((st = check_positive_finite(vs, "voxel_size"))) return st;
    if ((st = check_finite(ox, "off_x"))) return st;
    if ((st = check_finite(oy, "off_y"))) return st;
    if ((st = check_finite(oz, "off_z"))) return st;
    return check_step_resolution(step, make_guard_host(D, H, W, vs, ox, oy, oz, U, V, pix, step));
}

}  // namespace cbct

using namespace cbct;

extern "C" size_t cbct_fp_workspace_bytes(int64_t A, int64_t U, int64_t V) {
    if (A < 1 || U < 1 || V < 1) return 0;
    return align256((size_t)A * sizeof(ViewFrame)) + (size_t)A * fp_tiles(U, V) * 8 * sizeof(float);
}

static int fp_forward_impl(const float* volume, int64_t D, int64_t H, int64_t W, const float* const g[8],
                           int64_t A, float pixel_size, int64_t U, int64_t V, float step_size, float voxel_size,
                           float off_x, float off_y, float off_z, float* projections, const cbct_fp_options* opt,
                           void* workspace, size_t workspace_bytes, void* stream_) {
    int st;
    if ((st = check_fp_common(D, H, W, A, U, V, pixel_size, step_size, voxel_size, off_x, off_y, off_z))) return st;
    if ((st = check_geometry(g, A))) return st;
    CBCT_CHECK_ARG(volume != nullptr && projections != nullptr, "volume and projections must be non-null");
    FpCall call;
    if ((st = call_settings(opt, 0, call))) return st;
    CBCT_CHECK_ARG(workspace != nullptr && workspace_bytes >= cbct_fp_workspace_bytes(A, U, V),
                   "workspace too small for cbct_fp_forward");
    hipStream_t stream = (hipStream_t)stream_;
    ViewFrame* frames = (ViewFrame*)workspace;
    const GeomPtrs gp{g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7]};
    const FrameGuard guard = make_guard_host(D, H, W, voxel_size, off_x, off_y, off_z, U, V, pixel_size, step_size);
    if ((st = launch_frames(gp, A, frames, stream, &guard))) return st;
    const Box b = make_box_host(D, H, W, voxel_size, off_x, off_y, off_z);
    const int tu = (int)ceil_div(U, FP_TILE), tv = (int)ceil_div(V, FP_TILE);
    const int nblocks = (int)(A * tu * tv);
    const float inv_vs = 1.0f / voxel_size;
    const FpTuning& tf = call.t;
    if (tf.use_slab) {
        const SlabArgs args{volume, nullptr, frames, projections, nullptr, nullptr, b, (int)U, (int)V, tu, tv,
                            nblocks, pixel_size, step_size, voxel_size, inv_vs, tf.T, tf.cap,
                            tf.layout};
        return launch_fp_slab(0, !is_pow2(voxel_size), args, stream);
    }
    if (is_pow2(voxel_size))
        fp_forward_kernel<false><<<nblocks, FP_THREADS, 0, stream>>>(
            volume, b, frames, (int)U, (int)V, tu, tv, nblocks, pixel_size, step_size, voxel_size,
            inv_vs, projections);
    else
        fp_forward_kernel<true><<<nblocks, FP_THREADS, 0, stream>>>(
            volume, b, frames, (int)U, (int)V, tu, tv, nblocks, pixel_size, step_size, voxel_size,
            inv_vs, projections);
    CBCT_LAUNCHED("source-only forward kernel launch");
    return CBCT_OK;
}

extern "C" int cbct_fp_forward(const float* volume, int64_t D, int64_t H, int64_t W,
                               const float* angles, const float* sid, const float* sod,
                               const float* det_u0, const float* det_v0,
                               const float* src_x, const float* src_y, const float* src_z, int64_t A,
                               float pixel_size, int64_t U, int64_t V, float step_size,
                               float voxel_size, float off_x, float off_y, float off_z,
                               float* projections, void* workspace, size_t workspace_bytes,
                               void* stream_) {
    const float* g[8] = {angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z};
    return fp_forward_impl(volume, D, H, W, g, A, pixel_size, U, V, step_size, voxel_size, off_x, off_y, off_z,
                           projections, nullptr, workspace, workspace_bytes, stream_);
}

extern "C" int cbct_fp_forward_ex(const float* volume, int64_t D, int64_t H, int64_t W,
                                  const float* angles, const float* sid, const float* sod,
                                  const float* det_u0, const float* det_v0,
                                  const float* src_x, const float* src_y, const float* src_z, int64_t A,
                                  float pixel_size, int64_t U, int64_t V, float step_size,
                                  float voxel_size, float off_x, float off_y, float off_z,
                                  float* projections, const cbct_fp_options* options, void* workspace,
                                  size_t workspace_bytes, void* stream_) {
    const float* g[8] = {angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z};
    return fp_forward_impl(volume, D, H, W, g, A, pixel_size, U, V, step_size, voxel_size, off_x, off_y, off_z,
                           projections, options, workspace, workspace_bytes, stream_);
}

static size_t det_extra_bytes(int64_t D, int64_t H, int64_t W) {
    return align256((size_t)D * H * W * sizeof(unsigned long long)) + 256;
}

extern "C" size_t cbct_fp_backward_workspace_bytes_ex(int64_t D, int64_t H, int64_t W, int64_t A, int64_t U,
                                                      int64_t V, const cbct_fp_options* options) {
    if (D <= 0 || H <= 0 || W <= 0) return 0;
    FpCall call;
    if (call_settings(options, 1, call)) return 0;
    return cbct_fp_workspace_bytes(A, U, V) + (call.det ? det_extra_bytes(D, H, W) : 0);
}

extern "C" size_t cbct_fp_backward_workspace_bytes(int64_t D, int64_t H, int64_t W, int64_t A, int64_t U,
                                                   int64_t V) {
    return cbct_fp_backward_workspace_bytes_ex(D, H, W, A, U, V, nullptr);
}

extern "C" int cbct_fp_set_deterministic(int on) {
    CBCT_CHECK_ARG(on == 0 || on == 1, "deterministic must be 0 or 1");
    std::lock_guard<std::mutex> lock(g_defaults_mutex);
    g_deterministic = on;
    return CBCT_OK;
}

extern "C" int cbct_fp_get_deterministic(void) {
    std::lock_guard<std::mutex> lock(g_defaults_mutex);
    return g_deterministic;
}

static int fp_backward_impl(const float* grad_projection, const float* volume, int64_t D, int64_t H, int64_t W,
                            const float* const g[8], int64_t A, int64_t U, int64_t V, float pixel_size,
                            float step_size, float voxel_size, int compute_grad_volume, float off_x, float off_y,
                            float off_z, float* grad_volume, float* grad_geometry, bool geometry_optional,
                            const cbct_fp_options* opt, void* workspace, size_t workspace_bytes, void* stream_) {
    int st;
    if ((st = check_fp_common(D, H, W, A, U, V, pixel_size, step_size, voxel_size, off_x, off_y, off_z))) return st;
    if ((st = check_geometry(g, A))) return st;
    CBCT_CHECK_ARG(grad_projection != nullptr && volume != nullptr, "grad_projection and volume must be non-null");
    CBCT_CHECK_ARG(geometry_optional || grad_geometry != nullptr, "grad_geometry must be non-null");
    CBCT_CHECK_ARG(!compute_grad_volume || grad_volume != nullptr,
                   "grad_volume must be non-null when compute_grad_volume is set");
    const bool geometry = grad_geometry != nullptr;
    CBCT_CHECK_ARG(compute_grad_volume || geometry,
                   "nothing to compute: neither the volume gradient nor the geometry gradients are requested");
    FpCall call;
    if ((st = call_settings(opt, compute_grad_volume ? 1 : 0, call))) return st;
    const bool det = call.det && compute_grad_volume;
    const FpTuning& tb = call.t;
    CBCT_CHECK_ARG(!det || tb.use_slab, "the deterministic volume gradient needs the slab kernels");
    const size_t base_bytes = cbct_fp_workspace_bytes(A, U, V);
    CBCT_CHECK_ARG(workspace != nullptr &&
                   workspace_bytes >= base_bytes + (det ? det_extra_bytes(D, H, W) : 0),
                   "workspace too small for cbct_fp_backward (see cbct_fp_backward_workspace_bytes)");
    hipStream_t stream = (hipStream_t)stream_;
    ViewFrame* frames = (ViewFrame*)workspace;
    float* partials = (float*)((char*)workspace + align256((size_t)A * sizeof(ViewFrame)));
    const GeomPtrs gp{g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7]};
    const FrameGuard guard = make_guard_host(D, H, W, voxel_size, off_x, off_y, off_z, U, V, pixel_size, step_size);
    if ((st = launch_frames(gp, A, frames, stream, &guard))) return st;
    unsigned long long* gacc = nullptr;
    int* gexp = nullptr;
    const int64_t nvox = D * H * W;
    if (det) {
        gacc = (unsigned long long*)((char*)workspace + base_bytes);
        unsigned* maxbits = (unsigned*)((char*)gacc + align256((size_t)nvox * sizeof(unsigned long long)));
        gexp = (int*)(maxbits + 1);
        const unsigned zb = (unsigned)std::min<int64_t>(ceil_div(nvox, (int64_t)256), 16384);
        zero_u64_kernel<<<zb, 256, 0, stream>>>(gacc, nvox);
        zero_u64_kernel<<<1, 64, 0, stream>>>((unsigned long long*)maxbits, 1);
        CBCT_LAUNCHED("deterministic accumulator reset");
        const int64_t ng = A * U * V;
        if ((st = launch_absmax(grad_projection, ng, maxbits, stream))) return st;
        det_exponent_kernel<<<1, 1, 0, stream>>>(maxbits, step_size, gexp);
        CBCT_LAUNCHED("deterministic scale kernels");
    } else if (compute_grad_volume && (st = launch_fill(grad_volume, nvox, 0.0f, stream))) {
        return st;
    }
    const Box b = make_box_host(D, H, W, voxel_size, off_x, off_y, off_z);
    const int tu = (int)ceil_div(U, FP_TILE), tv = (int)ceil_div(V, FP_TILE);
    const int nblocks = (int)(A * tu * tv);
    const float inv_vs = 1.0f / voxel_size;
    if (tb.use_slab) {
        SlabArgs args{volume, grad_projection, frames, nullptr, compute_grad_volume ? grad_volume : nullptr,
                      partials, b, (int)U, (int)V, tu, tv, nblocks, pixel_size, step_size, voxel_size,
                      inv_vs, tb.T, tb.cap, tb.layout};
        args.gacc = gacc;
        args.gexp = gexp;
        if ((st = launch_fp_slab(compute_grad_volume ? 2 : 1, false, args, stream, geometry))) return st;
        if (det) {
            det_convert_kernel<<<(unsigned)ceil_div(nvox, (int64_t)256), 256, 0, stream>>>(gacc, nvox, gexp,
                                                                                         grad_volume);
            CBCT_LAUNCHED("deterministic volume gradient conversion");
        }
    } else if (compute_grad_volume) {
        fp_backward_kernel<true><<<nblocks, FP_THREADS, 0, stream>>>(
            grad_projection, volume, b, frames, (int)U, (int)V, tu, tv, nblocks, pixel_size,
            step_size, inv_vs, grad_volume, partials);
    } else {
        fp_backward_kernel<false><<<nblocks, FP_THREADS, 0, stream>>>(
            grad_projection, volume, b, frames, (int)U, (int)V, tu, tv, nblocks, pixel_size,
            step_size, inv_vs, nullptr, partials);
    }
    CBCT_LAUNCHED("source-only backward kernel launch");
    if (!geometry) return CBCT_OK;
    fp_reduce_geometry<<<(unsigned)A, 256, 0, stream>>>(partials, (int)A, tu * tv, grad_geometry);
    CBCT_LAUNCHED("geometry reduction kernel launch");
    return CBCT_OK;
}

extern "C" int cbct_fp_backward(const float* grad_projection, const float* volume,
                                int64_t D, int64_t H, int64_t W,
                                const float* angles, const float* sid, const float* sod,
                                const float* det_u0, const float* det_v0,
                                const float* src_x, const float* src_y, const float* src_z,
                                int64_t A, int64_t U, int64_t V,
                                float pixel_size, float step_size, float voxel_size,
                                int compute_grad_volume, float off_x, float off_y, float off_z,
                                float* grad_volume, float* grad_geometry,
                                void* workspace, size_t workspace_bytes, void* stream_) {
    const float* g[8] = {angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z};
    return fp_backward_impl(grad_projection, volume, D, H, W, g, A, U, V, pixel_size, step_size, voxel_size,
                            compute_grad_volume, off_x, off_y, off_z, grad_volume, grad_geometry, false, nullptr,
                            workspace, workspace_bytes, stream_);
}

extern "C" int cbct_fp_backward_ex(const float* grad_projection, const float* volume,
                                   int64_t D, int64_t H, int64_t W,
                                   const float* angles, const float* sid, const float* sod,
                                   const float* det_u0, const float* det_v0,
                                   const float* src_x, const float* src_y, const float* src_z,
                                   int64_t A, int64_t U, int64_t V,
                                   float pixel_size, float step_size, float voxel_size,
                                   int compute_grad_volume, float off_x, float off_y, float off_z,
                                   float* grad_volume, float* grad_geometry, const cbct_fp_options* options,
                                   void* workspace, size_t workspace_bytes, void* stream_) {
    const float* g[8] = {angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z};
    return fp_backward_impl(grad_projection, volume, D, H, W, g, A, U, V, pixel_size, step_size, voxel_size,
                            compute_grad_volume, off_x, off_y, off_z, grad_volume, grad_geometry, true, options,
                            workspace, workspace_bytes, stream_);
}

// ---- allocation-free adjoint
// The workspace of cbct_fp_backward_noalloc (view frames, per-tile geometry partials and, in
// deterministic mode, the int64 accumulator) lives in this fixed device array: each GPU the
// library runs on holds its own copy of it until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the symbol): the device address
// is looked up once per device with hipGetSymbolAddress and kept in g_scratch_ptr.
static __device__ __align__(256) unsigned char scratch[CBCT_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static void* g_scratch_ptr[MAX_DEVICES] = {};

// The current device's address of the scratch array (looked up on the first call on the device);
// the stream must belong to that device.
static int scratch_address(hipStream_t stream, void** out) {
    int dev = -1, st;
    if ((st = current_stream_device(stream, &dev))) return st;
    CBCT_CHECK_ARG(dev >= 0 && dev < MAX_DEVICES, "device ordinal beyond the library's 64-device table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        CBCT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch symbol address");
        CBCT_CHECK_ARG(p != nullptr, "scratch symbol address is null");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return CBCT_OK;
}

extern "C" int cbct_fp_backward_noalloc(const float* grad_projection, const float* volume,
                                        int64_t D, int64_t H, int64_t W,
                                        const float* angles, const float* sid, const float* sod,
                                        const float* det_u0, const float* det_v0,
                                        const float* src_x, const float* src_y, const float* src_z,
                                        int64_t A, int64_t U, int64_t V,
                                        float pixel_size, float step_size, float voxel_size,
                                        int compute_grad_volume, float off_x, float off_y, float off_z,
                                        float* grad_volume, float* grad_geometry, const cbct_fp_options* options,
                                        void* stream_) {
    int st;
    if ((st = check_fp_common(D, H, W, A, U, V, pixel_size, step_size, voxel_size, off_x, off_y, off_z))) return st;
    FpCall call;
    if ((st = call_settings(options, compute_grad_volume ? 1 : 0, call))) return st;
    const size_t need = cbct_fp_backward_workspace_bytes_ex(D, H, W, A, U, V, options);
    if (need > (size_t)CBCT_NOALLOC_SCRATCH_BYTES)
        return fail(CBCT_EWORKSPACE, "cbct_fp_backward_noalloc needs " + std::to_string(need) +
                                         " bytes of scratch; its fixed scratch holds " +
                                         std::to_string((size_t)CBCT_NOALLOC_SCRATCH_BYTES) +
                                         " (use cbct_fp_backward_ex with a workspace)");
    void* ws = nullptr;
    if ((st = scratch_address((hipStream_t)stream_, &ws))) return st;
    const float* g[8] = {angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z};
    return fp_backward_impl(grad_projection, volume, D, H, W, g, A, U, V, pixel_size, step_size, voxel_size,
                            compute_grad_volume, off_x, off_y, off_z, grad_volume, grad_geometry, true, options, ws,
                            (size_t)CBCT_NOALLOC_SCRATCH_BYTES, stream_);
}

extern "C" int cbct_fp_count_samples(int64_t D, int64_t H, int64_t W,
                                     const float* angles, const float* sid, const float* sod,
                                     const float* det_u0, const float* det_v0,
                                     const float* src_x, const float* src_y, const float* src_z, int64_t A,
                                     float pixel_size, int64_t U, int64_t V, float step_size,
                                     float voxel_size, float off_x, float off_y, float off_z,
                                     uint64_t* count, void* workspace, size_t workspace_bytes,
                                     void* stream_) {
    int st;
    if ((st = check_fp_common(D, H, W, A, U, V, pixel_size, step_size, voxel_size, off_x, off_y, off_z))) return st;
    const float* g[8] = {angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z};
    if ((st = check_geometry(g, A))) return st;
    CBCT_CHECK_ARG(count != nullptr, "count must be non-null");
    CBCT_CHECK_ARG(workspace != nullptr && workspace_bytes >= cbct_fp_workspace_bytes(A, U, V),
                   "workspace too small for cbct_fp_count_samples");
    hipStream_t stream = (hipStream_t)stream_;
    ViewFrame* frames = (ViewFrame*)workspace;
    const GeomPtrs gp{angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z};
    const FrameGuard guard = make_guard_host(D, H, W, voxel_size, off_x, off_y, off_z, U, V, pixel_size, step_size);
    if ((st = launch_frames(gp, A, frames, stream, &guard))) return st;
    const Box b = make_box_host(D, H, W, voxel_size, off_x, off_y, off_z);
    const int tu = (int)ceil_div(U, FP_TILE), tv = (int)ceil_div(V, FP_TILE);
    const int nblocks = (int)(A * tu * tv);
    fp_count_kernel<<<nblocks, FP_THREADS, 0, stream>>>(b, frames, (int)U, (int)V, tu, tv, nblocks,
                                                        pixel_size, step_size,
                                                        (unsigned long long*)count);
    CBCT_LAUNCHED("sample count kernel launch");
    return CBCT_OK;
}

extern "C" int cbct_fp_set_tuning_op(int op, int use_slab, int slab_thickness, int lds_floats, int layout) {
    CBCT_CHECK_ARG(op == 0 || op == 1, "op must be 0 (gathering kernels) or 1 (volume-gradient adjoint)");
    const cbct_fp_options o{use_slab, slab_thickness, lds_floats, layout, 0};
    const int st = check_options(o);
    if (st) return st;
    std::lock_guard<std::mutex> lock(g_defaults_mutex);
    g_tuning_op[op] = FpTuning{use_slab, slab_thickness, lds_floats, layout};
    return CBCT_OK;
}

extern "C" int cbct_fp_get_tuning_op(int op, int* out4) {
    CBCT_CHECK_ARG(op == 0 || op == 1, "op must be 0 (gathering kernels) or 1 (volume-gradient adjoint)");
    CBCT_CHECK_ARG(out4 != nullptr, "out4 must be non-null");
    std::lock_guard<std::mutex> lock(g_defaults_mutex);
    const FpTuning& t = g_tuning_op[op];
    out4[0] = t.use_slab;
    out4[1] = t.T;
    out4[2] = t.cap;
    out4[3] = t.layout;
    return CBCT_OK;
}

extern "C" int cbct_fp_set_tuning(int use_slab, int slab_thickness, int lds_floats, int layout) {
    int st = cbct_fp_set_tuning_op(0, use_slab, slab_thickness, lds_floats, layout);
    if (st == CBCT_OK) st = cbct_fp_set_tuning_op(1, use_slab, slab_thickness, lds_floats, layout);
    return st;
}

extern "C" int cbct_fp_get_tuning(int* out4) { return cbct_fp_get_tuning_op(0, out4); }

extern "C" int cbct_fp_default_options(int op, cbct_fp_options* out) {
    CBCT_CHECK_ARG(op == 0 || op == 1, "op must be 0 (gathering kernels) or 1 (volume-gradient adjoint)");
    CBCT_CHECK_ARG(out != nullptr, "out must be non-null");
    FpCall c;
    call_settings(nullptr, op, c);
    *out = cbct_fp_options{c.t.use_slab, c.t.T, c.t.cap, c.t.layout, c.det};
    return CBCT_OK;
}

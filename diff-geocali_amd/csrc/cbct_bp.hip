// Weighted voxel-driven backprojector and its backward for gfx950.
//
// Work assignment (forward): one wave owns one (x, y) voxel column and 64*K voxels of it
// along z (lane = z, K voxels per lane); the backward groups 8x8 columns per workgroup
// (see bp_backward_kernel).  Along a column only the detector row
// v changes, and v is the contiguous axis of the [A,U,V] projections, so the
// four bilinear taps of a wave fall in a few cache lines.  All view-dependent
// terms that do not depend on z (rotation, magnification t, u index, weight)
// are evaluated once per lane and view and reused for its K voxels.  The
// per-view constants are wave-uniform and read through the scalar cache.
//
// Arithmetic follows csrc/src_cbct_backprojector_cuda.cu:37-95 (forward) and
// :132-292 (backward); the detector coordinates use the reciprocal pixel size (one FMA instead of
// a division per voxel-view), which differs from the reference's division by FP32 rounding.
#include "cbct_common.h"

#include <algorithm>

namespace cbct {

constexpr int BP_WAVES = 4;  // waves (columns) per workgroup

struct BpGeom {
    int U, V, D, H, W, A;
    float pix, inv_pix, vs, offx, offy, offz, hx, hy, hz;
};

// out = scale * sum over chunks of (sum over the chunk's views), clamped at 0
// when clamp is set.  chunk = A gives the plain backprojection.
template <int K>
__global__ __launch_bounds__(64 * BP_WAVES) void bp_forward_kernel(
    const float* __restrict__ proj, const ViewFrame* __restrict__ frames, BpGeom g,
    int chunk, float scale, int clamp, float* __restrict__ vol) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = __builtin_amdgcn_readfirstlane(blockIdx.x * BP_WAVES + wave);
    if (col >= g.H * g.W) return;
    const int x = col % g.W, y = col / g.W;
    const int zb = blockIdx.y * 64 * K + lane;
    const float vx = fmaf((float)x - g.hx, g.vs, g.offx);
    const float vy = fmaf((float)y - g.hy, g.vs, g.offy);
    float vz[K], total[K], acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        vz[k] = fmaf((float)(zb + 64 * k) - g.hz, g.vs, g.offz);
        total[k] = 0.f;
        acc[k] = 0.f;
    }
    const float Um1 = (float)g.U - 1.0f, Vm1 = (float)g.V - 1.0f;
    int next = chunk;
    for (int a = 0; a < g.A; ++a) {
        if (a == next) {
#pragma unroll
            for (int k = 0; k < K; ++k) { total[k] += acc[k]; acc[k] = 0.f; }
            next += chunk;
        }
        const ViewFrame f = frames[a];
        const float px = fmaf(vx, f.c, vy * f.s);
        const float py = fmaf(-vx, f.s, vy * f.c);
        const float m = f.sod + f.dsn;
        const float q = m - py;
        const float l = f.sid + f.dsn;
        if (q <= 1e-6f || l <= 1e-6f) continue;
        const float iq = 1.0f / q;
        const float t = l * iq;
        const float uph = fmaf(t, px - f.dsu, f.dsu);
        const float sw = m * iq;
        const float gw = sw * sw;
        const float ui = fmaf(uph, g.inv_pix, f.u0);
        if (!(ui >= 0.0f && ui < Um1)) continue;
        const int iu = (int)floorf(ui);
        const float fu = ui - (float)iu, mu = 1.0f - fu;
        const float* row0 = proj + ((int64_t)a * g.U + iu) * g.V;
        const float* row1 = row0 + g.V;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const float vph = fmaf(t, vz[k] - f.dsv, f.dsv);
            const float vi = fmaf(vph, g.inv_pix, f.v0);
            if (vi >= 0.0f && vi < Vm1) {
                // vi >= 0: truncation is the floor, and the fraction is exact; an unsigned 32-bit
                // byte offset lets the loads take the wave-uniform row address as their base
                const int iv = (int)vi;
                const float fv = __builtin_amdgcn_fractf(vi), mv = 1.0f - fv;
                const unsigned bo = 4u * (unsigned)iv;      // byte offset in the row (< 2^31)
                const float* const q0 = reinterpret_cast<const float*>(reinterpret_cast<const char*>(row0) + bo);
                const float* const q1 = reinterpret_cast<const float*>(reinterpret_cast<const char*>(row1) + bo);
                const float p00 = q0[0], p01 = q0[1], p10 = q1[0], p11 = q1[1];
                const float interp = fmaf(fu * fv, p11, fmaf(mu * fv, p01, fmaf(fu * mv, p10, mu * mv * p00)));
                acc[k] = fmaf(gw, interp, acc[k]);
            }
        }
    }
    const int64_t plane = (int64_t)g.H * g.W;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int z = zb + 64 * k;
        if (z < g.D) {
            float r = (total[k] + acc[k]) * scale;
            if (clamp) r = fmaxf(r, 0.0f);
            vol[z * plane + (int64_t)y * g.W + x] = r;
        }
    }
}

// Backward: per voxel-view, the bilinear transpose scatters into gproj and the
// analytic derivatives of weight*interp feed the eight geometry sums, which
// are reduced over the wave and added once per wave and view.
// Backward: the transpose of the bilinear interpolation (the projection gradient) and the
// eight per-view geometry derivatives (csrc/src_cbct_backprojector_cuda.cu:101-295).
//
// One workgroup owns an 8x8 block of voxel columns and 64*K voxels of each along z; each wave
// takes 16 of the block's columns (lane = z).  Per view the z-independent factors of a column
// (its detector row and weight, and the coefficients that turn its four per-voxel sums into the
// eight geometry terms) are computed once, by lane L for the block's column L, and handed to
// the wave that takes that column through LDS (16-byte broadcast reads), instead of by all 64
// lanes of that wave for each of its 16 columns: 133 -> 88 VALU instructions per voxel-view,
// 18.1 -> 15.4 ms at 256^3 x 360.  Per view the block's taps fall in a small
// detector window (columns of the block share a few detector rows u, and a column covers a run
// of v): the workgroup bounds that window exactly (same float expressions as the taps), clears
// an LDS tile for it, adds its projection-gradient contributions there as 32-bit fixed-point
// integers (ds_add_u32; float LDS atomics are ~40x slower on gfx950), and flushes the tile with
// one coalesced float atomic per touched detector cell -- about an order of magnitude fewer
// global atomics than four per voxel-view.  A window larger than the tile falls back to direct
// global atomics for that view.  The fixed-point scale 2^e per view keeps the largest possible
// cell total (max|grad| x max weight x columns x voxels of a column per cell) below 2^30.
//
// The eight geometry sums of a view are combined over the workgroup's waves in a fixed order
// and stored as the workgroup's partial (no atomics); launch_reduce_geometry sums the partials
// of each view in a fixed order, so grad_geometry is bitwise reproducible.  With DET the
// projection gradient is accumulated as int64 fixed point at a per-view scale 2^E[a] (E from
// max|grad_volume| x the view's largest weight, so one sample stays below 2^40 units): the LDS
// cells are converted exactly from 2^e to 2^E[a] (rounded when e > E[a]) and added with int64
// atomics, whose sum does not depend on their order.
constexpr int BT = 8;                          // column block edge
constexpr int BT_COLS_PER_WAVE = BT * BT / BP_WAVES;
constexpr int BT_CAP = 8192;                   // LDS tile cells
constexpr int BP_COLF = 24;                    // per-column factor floats (22 used)

struct BpColumn {
    float px, py, m, q, l, iq, iq2, iq3, t, du, gw, ui;
    bool ok;
};

__device__ __forceinline__ BpColumn bp_column(const ViewFrame& f, float vx, float vy, float ip) {
    BpColumn c;
    c.px = fmaf(vx, f.c, vy * f.s);
    c.py = fmaf(-vx, f.s, vy * f.c);
    c.m = f.sod + f.dsn;
    c.q = c.m - c.py;
    c.l = f.sid + f.dsn;
    c.ok = !(c.q <= 1e-6f || c.l <= 1e-6f);
    c.iq = 1.0f / c.q;
    c.iq2 = c.iq * c.iq;
    c.iq3 = c.iq2 * c.iq;
    c.t = c.l * c.iq;
    c.du = c.px - f.dsu;
    const float uph = fmaf(c.t, c.du, f.dsu);
    c.gw = (c.m * c.m) * c.iq2;
    c.ui = fmaf(uph, ip, f.u0);
    return c;
}

__device__ __forceinline__ float bp_vi(const ViewFrame& f, float t, float vz, float ip) {
    const float dv = vz - f.dsv;
    const float vph = fmaf(t, dv, f.dsv);
    return fmaf(vph, ip, f.v0);
}

// Wave-wide (64 lanes) reductions through DPP row operations: VALU only, where the __shfl_xor
// butterfly costs an LDS permute per step.  The two quad permutations, the half-row and the row
// mirror leave every lane of a row of 16 with that row's total; the row broadcasts of lanes 15 and
// 31 carry rows 0..2 into row 3, and lane 63's value is returned as a wave-uniform scalar.  Every
// lane of the wave must be active.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ float dpp_from(float v, float ident) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(ident), __float_as_int(v), CTRL, ROWS, 0xF, false));
}
template <class Op>
__device__ __forceinline__ float wave_reduce(float v, float ident, Op op) {
    v = op(v, dpp_from<0xB1>(v, ident));         // quad_perm [1,0,3,2]
    v = op(v, dpp_from<0x4E>(v, ident));         // quad_perm [2,3,0,1]
    v = op(v, dpp_from<0x141>(v, ident));        // row_half_mirror
    v = op(v, dpp_from<0x140>(v, ident));        // row_mirror
    v = op(v, dpp_from<0x142, 0xA>(v, ident));   // row_bcast:15 into rows 1 and 3
    v = op(v, dpp_from<0x143, 0xC>(v, ident));   // row_bcast:31 into rows 2 and 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ float wave_min(float v) {
    return wave_reduce(v, __builtin_huge_valf(), [](float a, float b) { return fminf(a, b); });
}
__device__ __forceinline__ float wave_max(float v) {
    return wave_reduce(v, -__builtin_huge_valf(), [](float a, float b) { return fmaxf(a, b); });
}
__device__ __forceinline__ float wave_add(float v) {
    return wave_reduce(v, 0.0f, [](float a, float b) { return a + b; });
}

template <int K, bool DET>
__global__ __launch_bounds__(64 * BP_WAVES) void bp_backward_kernel(
    const float* __restrict__ gvol, const float* __restrict__ proj,
    const ViewFrame* __restrict__ frames, BpGeom g,
    float* __restrict__ gproj, float* __restrict__ partials,
    unsigned long long* __restrict__ gacc, const int* __restrict__ gexp) {
    __shared__ unsigned tile[BT_CAP];
    __shared__ float red[BP_WAVES];
    __shared__ float gred[BP_WAVES][8];
    // per view, the z-independent factors of each of the block's 64 columns (BP_COLF floats each)
    __shared__ float4 colf[BT * BT * BP_COLF / 4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int blk = blockIdx.y * gridDim.x + blockIdx.x, nblk = gridDim.x * gridDim.y;
    const int nbx = (g.W + BT - 1) / BT;
    const int bx = blockIdx.x % nbx, by = blockIdx.x / nbx;
    const int z0 = blockIdx.y * 64 * K;
    const int64_t plane = (int64_t)g.H * g.W;
    float vz[K];
#pragma unroll
    for (int k = 0; k < K; ++k) vz[k] = fmaf((float)(z0 + lane + 64 * k) - g.hz, g.vs, g.offz);
    // the wave's columns: block column c = wave + BP_WAVES * j
    float gv[BT_COLS_PER_WAVE][K];
    float amax = 0.f;
#pragma unroll
    for (int j = 0; j < BT_COLS_PER_WAVE; ++j) {
        const int c = wave + BP_WAVES * j;
        const int x = bx * BT + c % BT, y = by * BT + c / BT;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int z = z0 + lane + 64 * k;
            gv[j][k] = (x < g.W && y < g.H && z < g.D) ? gvol[z * plane + (int64_t)y * g.W + x] : 0.0f;
            amax = fmaxf(amax, fabsf(gv[j][k]));
        }
    }
    for (int s = 32; s >= 1; s >>= 1) amax = fmaxf(amax, __shfl_xor(amax, s, 64));
    if (lane == 0) red[wave] = amax;
    __syncthreads();
    amax = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    // the block column of this lane for the window bounds, and the chunk's z extremes
    const int bcx = bx * BT + lane % BT, bcy = by * BT + lane / BT;
    const bool bcol = bcx < g.W && bcy < g.H;
    const float bvx = fmaf((float)bcx - g.hx, g.vs, g.offx), bvy = fmaf((float)bcy - g.hy, g.vs, g.offy);
    const int zhi = min(z0 + 64 * K, g.D) - 1;
    const float vz_lo = fmaf((float)z0 - g.hz, g.vs, g.offz), vz_hi = fmaf((float)zhi - g.hz, g.vs, g.offz);
    const float Um1 = (float)g.U - 1.0f, Vm1 = (float)g.V - 1.0f;
    const float ip = g.inv_pix;
    const float INF = __builtin_huge_valf();

    for (int a = 0; a < g.A; ++a) {
        const ViewFrame f = frames[a];
        // ---- exact detector window of the block's taps (uniform over the workgroup)
        float umin = INF, umax = -INF, vmin = INF, vmax = -INF, gwmax = 0.f, tmin = INF;
        BpColumn c{};
        if (bcol) {
            c = bp_column(f, bvx, bvy, ip);
            if (c.ok) {
                const float v1 = bp_vi(f, c.t, vz_lo, ip), v2 = bp_vi(f, c.t, vz_hi, ip);
                umin = umax = c.ui;
                vmin = fminf(v1, v2);
                vmax = fmaxf(v1, v2);
                gwmax = c.gw;
                tmin = c.t;
            }
        }
        umin = wave_min(umin); umax = wave_max(umax);
        vmin = wave_min(vmin); vmax = wave_max(vmax);
        gwmax = wave_max(gwmax); tmin = wave_min(tmin);
        float* part = partials + ((int64_t)a * nblk + blk) * 8;
        if (!(umax >= 0.f && umin < Um1 && vmax >= 0.f && vmin < Vm1)) {      // no tap on the detector
            if (threadIdx.x < 8) part[threadIdx.x] = 0.f;
            continue;
        }
        const int Ea = DET ? gexp[a] : 0;
        const int u_lo = (int)fmaxf(floorf(umin), 0.f), u_hi = (int)fminf(floorf(umax), (float)(g.U - 2));
        const int v_lo = (int)fmaxf(floorf(vmin), 0.f), v_hi = (int)fminf(floorf(vmax), (float)(g.V - 2));
        const int nu = u_hi - u_lo + 2, nv = v_hi - v_lo + 2;
        const bool staged = nu * nv <= BT_CAP;
        float scale = 1.f, inv_scale = 1.f;
        int e = 0;
        if (staged) {
            const float spacing = g.vs * tmin * ip;                  // detector rows per voxel along z
            const float nz = spacing > 0.f ? fminf(floorf(2.0f / spacing) + 2.0f, 64.0f * K) : 64.0f * K;
            const float bound = amax * gwmax * (float)(BT * BT) * nz;
            if (bound > 0.f) {
                int k2;
                frexpf(1073741824.f / bound, &k2);
                e = max(-100, min(100, k2 - 1));
            }
            scale = ldexpf(1.f, e);
            inv_scale = ldexpf(1.f, -e);
            for (int i = threadIdx.x; i < nu * nv; i += 64 * BP_WAVES) tile[i] = 0u;
        }
        // The z-independent factors of lane L's column (the block's column L), once per view, for
        // the wave that takes that column (L mod BP_WAVES): the same expressions as the per-column
        // code computed them in every lane of that wave for every column.
        if ((lane & (BP_WAVES - 1)) == wave) {
            const bool cvalid = bcol && c.ok && c.ui >= 0.0f && c.ui < Um1;
            float cf[BP_COLF];
#pragma unroll
            for (int i = 0; i < BP_COLF; ++i) cf[i] = 0.f;
            if (cvalid) {
                const float iq = c.iq, iq2 = c.iq2, iq3 = c.iq3, t = c.t, du = c.du;
                const float m = c.m, l = c.l, q = c.q, px = c.px, py = c.py;
                const float m2 = m * m;
                const int iu = (int)floorf(c.ui);
                const float fu = c.ui - (float)iu;
                const float dt_sod = -l * iq2;
                const float dW_m = 2.0f * m * iq2 - 2.0f * m2 * iq3;     // d(m^2/q^2)/dm at fixed q-m
                const float dsu_b = f.dsn, dsn_b = -f.dsu;
                const float dq_b = dsn_b + px;
                const float dt_b = (dsn_b * q - l * dq_b) * iq2;
                const float dU_b = fmaf(du, dt_b, fmaf(1.0f - t, dsu_b, t * py));
                const float dW_b = 2.0f * m * iq2 * dsn_b - 2.0f * m2 * iq3 * dq_b;
                const float qml = q - l;
                const float dt_sx = -f.s * qml * iq2;
                const float dU_sx = fmaf(du, dt_sx, (1.0f - t) * f.c);
                const float dW_sx = dW_m * -f.s;
                const float dt_sy = f.c * qml * iq2;
                const float dU_sy = fmaf(du, dt_sy, (1.0f - t) * f.s);
                const float dW_sy = dW_m * f.c;
                cf[0] = __int_as_float(1); cf[1] = __int_as_float(iu); cf[2] = fu; cf[3] = 1.0f - fu;
                cf[4] = t; cf[5] = c.gw; cf[6] = __int_as_float((iu - u_lo) * nv);
                cf[7] = iq * ip; cf[8] = du * iq * ip;                        // sid
                cf[9] = dt_sod * ip; cf[10] = du * dt_sod * ip; cf[11] = dW_m;  // sod
                cf[12] = dt_b * ip; cf[13] = dU_b * ip; cf[14] = dW_b;        // beta
                cf[15] = dt_sx * ip; cf[16] = dU_sx * ip; cf[17] = dW_sx;     // source x
                cf[18] = dt_sy * ip; cf[19] = dU_sy * ip; cf[20] = dW_sy;     // source y
                cf[21] = (1.0f - t) * ip;                                     // source z
            }
#pragma unroll
            for (int i = 0; i < BP_COLF / 4; ++i)
                colf[lane * (BP_COLF / 4) + i] = make_float4(cf[4 * i], cf[4 * i + 1], cf[4 * i + 2], cf[4 * i + 3]);
        }
        __syncthreads();
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f, s5 = 0.f, s6 = 0.f, s7 = 0.f;
#pragma unroll 1
        for (int j = 0; j < BT_COLS_PER_WAVE; ++j) {
            const int cidx = wave + BP_WAVES * j;
            // (six 16-byte reads, all issued before the column's validity is looked at: volatile
            // keeps each one whole)
            typedef float f4 __attribute__((ext_vector_type(4)));
            f4 c4[BP_COLF / 4];
#pragma unroll
            for (int i = 0; i < BP_COLF / 4; ++i)
                c4[i] = *(const volatile __attribute__((address_space(3))) f4*)(colf + cidx * (BP_COLF / 4) + i);
            const float cf[BP_COLF] = {c4[0].x, c4[0].y, c4[0].z, c4[0].w, c4[1].x, c4[1].y, c4[1].z, c4[1].w,
                                       c4[2].x, c4[2].y, c4[2].z, c4[2].w, c4[3].x, c4[3].y, c4[3].z, c4[3].w,
                                       c4[4].x, c4[4].y, c4[4].z, c4[4].w, c4[5].x, c4[5].y, c4[5].z, c4[5].w};
            if (__float_as_int(cf[0]) == 0) continue;   // column outside the volume, behind the source or off the detector
            const int iu = __float_as_int(cf[1]);
            const float fu = cf[2], mu = cf[3], t = cf[4], gw = cf[5];
            const int row = __float_as_int(cf[6]);
            // the tap offsets fit in 32 bits: A * U * V <= INT32_MAX (host check); the column's
            // row is wave-uniform, so its address goes to scalar registers and each tap load
            // adds an unsigned 32-bit lane offset to it
            const int base = __builtin_amdgcn_readfirstlane((a * g.U + iu) * g.V);
            const float* const row0 = proj + base;
            const float* const row1 = row0 + g.V;
            // Every geometry term of a voxel is a combination, with coefficients constant along the
            // column, of four per-voxel products: gb*Iv*dv, gb*Iu, gb*Iv and g*interp.  The column
            // sums those four and combines them once (the same terms as :214-292, summed in
            // another order).
            float P1 = 0.f, P2 = 0.f, P3 = 0.f, P4 = 0.f;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const float dv = vz[k] - f.dsv;
                const float vph = fmaf(t, dv, f.dsv);
                const float vi = fmaf(vph, ip, f.v0);
                if (z0 + lane + 64 * k < g.D && vi >= 0.0f && vi < Vm1) {
                    // vi >= 0: truncation is the floor, and the fraction is exact
                    const int iv = (int)vi;
                    const float fv = __builtin_amdgcn_fractf(vi), mv = 1.0f - fv;
                    const int k00 = base + iv;
                    const unsigned bo = 4u * (unsigned)iv;  // byte offset in the row (< 2^31)
                    const float* const q0 = reinterpret_cast<const float*>(reinterpret_cast<const char*>(row0) + bo);
                    const float* const q1 = reinterpret_cast<const float*>(reinterpret_cast<const char*>(row1) + bo);
                    const float p00 = q0[0], p01 = q0[1];
                    const float p10 = q1[0], p11 = q1[1];
                    const float interp = fmaf(fu * fv, p11, fmaf(mu * fv, p01, fmaf(fu * mv, p10, mu * mv * p00)));
                    const float Iu = fmaf(fv, p11 - p01, mv * (p10 - p00));
                    const float Iv = fmaf(fu, p11 - p10, mu * (p01 - p00));
                    const float gb = gv[j][k] * gw;
                    if (staged) {
                        // scale is a power of two, so (gb scale) w1 w2 rounds exactly as (gb w1 w2) scale
                        const float gbs = gb * scale, gm = gbs * mu, gf = gbs * fu;
                        unsigned* q0 = tile + row + (iv - v_lo);
                        atomicAdd(q0, (unsigned)__float2int_rn(gm * mv));
                        atomicAdd(q0 + nv, (unsigned)__float2int_rn(gf * mv));
                        atomicAdd(q0 + 1, (unsigned)__float2int_rn(gm * fv));
                        atomicAdd(q0 + nv + 1, (unsigned)__float2int_rn(gf * fv));
                    } else if (DET) {
                        atomicAdd(gacc + k00, (unsigned long long)__double2ll_rn(ldexp((double)(gb * mu * mv), Ea)));
                        atomicAdd(gacc + k00 + g.V, (unsigned long long)__double2ll_rn(ldexp((double)(gb * fu * mv), Ea)));
                        atomicAdd(gacc + k00 + 1, (unsigned long long)__double2ll_rn(ldexp((double)(gb * mu * fv), Ea)));
                        atomicAdd(gacc + k00 + g.V + 1,
                                  (unsigned long long)__double2ll_rn(ldexp((double)(gb * fu * fv), Ea)));
                    } else {
                        atomicAdd(gproj + k00, gb * mu * mv);
                        atomicAdd(gproj + k00 + g.V, gb * fu * mv);
                        atomicAdd(gproj + k00 + 1, gb * mu * fv);
                        atomicAdd(gproj + k00 + g.V + 1, gb * fu * fv);
                    }
                    const float gbIv = gb * Iv;
                    P1 = fmaf(gbIv, dv, P1);
                    P2 = fmaf(gb, Iu, P2);
                    P3 += gbIv;
                    P4 = fmaf(gv[j][k], interp, P4);
                }
            }
            // sid (:214-219): gb (Iv dv + Iu du) / q / pix
            s1 += fmaf(P1, cf[7], P2 * cf[8]);
            // sod (:222-230)
            s2 += fmaf(P1, cf[9], fmaf(P2, cf[10], P4 * cf[11]));
            // beta (:233-250)
            s0 += fmaf(P1, cf[12], fmaf(P2, cf[13], P4 * cf[14]));
            // detector centres (:252-253)
            s3 += P2;
            s4 += P3;
            // source offsets (:255-292)
            s5 += fmaf(P1, cf[15], fmaf(P2, cf[16], P4 * cf[17]));
            s6 += fmaf(P1, cf[18], fmaf(P2, cf[19], P4 * cf[20]));
            s7 += P3 * cf[21];
        }
        s0 = wave_add(s0); s1 = wave_add(s1); s2 = wave_add(s2); s3 = wave_add(s3);
        s4 = wave_add(s4); s5 = wave_add(s5); s6 = wave_add(s6); s7 = wave_add(s7);
        if (lane == 0) {
            gred[wave][0] = s0; gred[wave][1] = s1; gred[wave][2] = s2; gred[wave][3] = s3;
            gred[wave][4] = s4; gred[wave][5] = s5; gred[wave][6] = s6; gred[wave][7] = s7;
        }
        __syncthreads();
        if (threadIdx.x < 8) {
            const int k = threadIdx.x;
            part[k] = (gred[0][k] + gred[1][k]) + (gred[2][k] + gred[3][k]);
        }
        if (staged) {
            // the tile's cells i = r * nv + c, r < nu, walked with every lane busy: from i to
            // i + 256 the column moves by 256 mod nv and the row by 256 / nv or one more
            const int n = nu * nv, dq = (64 * BP_WAVES) / nv, dr = (64 * BP_WAVES) - dq * nv;
            int r = threadIdx.x / nv, cc = threadIdx.x - r * nv;
            const int64_t row0 = ((int64_t)a * g.U + u_lo) * g.V + v_lo;
            const int sh = Ea - e;      // (DET) cell units 2^-e -> accumulator units 2^-Ea
            for (int i = threadIdx.x; i < n; i += 64 * BP_WAVES) {
                const int iv = (int)tile[i];
                if (iv != 0) {
                    const int64_t k = row0 + (int64_t)r * g.V + cc;
                    if (DET) atomicAdd(gacc + k, (unsigned long long)__double2ll_rn(ldexp((double)iv, sh)));
                    else atomicAdd(gproj + k, (float)iv * inv_scale);
                }
                cc += dr;
                r += dq;
                if (cc >= nv) { cc -= nv; ++r; }
            }
        }
        __syncthreads();
    }
}

static_assert(BP_WAVES == 4, "bp_backward_kernel combines the geometry sums of four waves");

// Per-view accumulator exponent of the deterministic backward: E[a] = 40 - k where
// max|grad_volume| x (largest weight of a column whose u lands on the detector) = f 2^k,
// f in [0.5, 1), so that one sample's contribution stays below 2^40 units of 2^-E[a].
__global__ __launch_bounds__(256) void bp_det_exponent_kernel(const ViewFrame* __restrict__ frames, BpGeom g,
                                                              const unsigned* __restrict__ maxbits,
                                                              int* __restrict__ gexp) {
    __shared__ float red[4];
    const int a = blockIdx.x;
    const ViewFrame f = frames[a];
    const float Um1 = (float)g.U - 1.0f;
    float wmax = 0.f;
    for (int i = threadIdx.x; i < g.H * g.W; i += 256) {
        const int x = i % g.W, y = i / g.W;
        const BpColumn c = bp_column(f, fmaf((float)x - g.hx, g.vs, g.offx), fmaf((float)y - g.hy, g.vs, g.offy),
                                     g.inv_pix);
        if (c.ok && c.ui >= 0.0f && c.ui < Um1) wmax = fmaxf(wmax, c.gw);
    }
    for (int s = 32; s >= 1; s >>= 1) wmax = fmaxf(wmax, __shfl_xor(wmax, s, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = wmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float m = __uint_as_float(maxbits[0]) * fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        int E = 0;
        if (m > 0.f && isfinite(m)) {
            int k;
            frexpf(m, &k);
            E = max(-60, min(120, 40 - k));
        }
        gexp[a] = E;
    }
}

__global__ void bp_det_convert_kernel(const unsigned long long* __restrict__ acc, int64_t n, int64_t per_view,
                                      const int* __restrict__ gexp, float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = (float)ldexp((double)(long long)acc[i], -gexp[i / per_view]);
}

static int check_bp_common(int64_t A, int64_t U, int64_t V, int64_t D, int64_t H, int64_t W,
                           float pix, float vs, float ox, float oy, float oz) {
    CBCT_CHECK_ARG(A >= 1 && U >= 2 && V >= 2, "proj must have shape [A,U,V] with A >= 1 and U,V >= 2");
    CBCT_CHECK_ARG(A * U * V <= INT32_MAX, "proj exceeds the kernel's 32-bit indexing limit");
    CBCT_CHECK_ARG(D > 0 && H > 0 && W > 0 && D <= INT32_MAX && H <= INT32_MAX && W <= INT32_MAX,
                   "D,H,W must be positive and fit in a 32-bit integer");
    CBCT_CHECK_ARG(D <= INT32_MAX / H && D * H <= INT32_MAX / W, "volume exceeds the kernel's 32-bit indexing limit");
    // z is covered by workgroups of 64 voxels (or more) along the launch grid's second dimension
    CBCT_CHECK_ARG(D <= 64 * 65535, "D exceeds the backprojector's launch limit of 4194240 voxels");
    int st;
    if ((st = check_positive_finite(pix, "pixel_size"))) return st;
    if ((st = check_positive_finite(vs, "voxel_size"))) return st;
    if ((st = check_finite(ox, "off_x"))) return st;
    if ((st = check_finite(oy, "off_y"))) return st;
    if ((st = check_finite(oz, "off_z"))) return st;
    return CBCT_OK;
}

static BpGeom make_bp_geom(int64_t A, int64_t U, int64_t V, int64_t D, int64_t H, int64_t W,
                           float pix, float vs, float ox, float oy, float oz) {
    BpGeom g;
    g.A = (int)A; g.U = (int)U; g.V = (int)V; g.D = (int)D; g.H = (int)H; g.W = (int)W;
    g.pix = pix; g.inv_pix = 1.0f / pix; g.vs = vs;
    g.offx = ox; g.offy = oy; g.offz = oz;
    g.hx = (float)W * 0.5f; g.hy = (float)H * 0.5f; g.hz = (float)D * 0.5f;
    return g;
}

static int bp_k(int64_t D) { return D > 128 ? 4 : (D > 64 ? 2 : 1); }

// Backward launch shape: 8x8 columns x 64*K voxels per workgroup.
static int bp_bwd_k(int64_t D) { return bp_k(D) >= 2 ? 2 : 1; }
static dim3 bp_bwd_grid(int64_t D, int64_t H, int64_t W) {
    return dim3((unsigned)(ceil_div(W, (int64_t)BT) * ceil_div(H, (int64_t)BT)),
                (unsigned)ceil_div(D, 64 * bp_bwd_k(D)));
}
static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

int launch_bp_forward(const float* proj, const ViewFrame* frames, const BpGeom& g, int chunk,
                      float scale, int clamp, float* vol, hipStream_t stream) {
    const int K = bp_k(g.D);
    dim3 grid((unsigned)ceil_div((int64_t)g.H * g.W, BP_WAVES), (unsigned)ceil_div(g.D, 64 * K));
    if (K == 4)
        bp_forward_kernel<4><<<grid, 64 * BP_WAVES, 0, stream>>>(proj, frames, g, chunk, scale, clamp, vol);
    else if (K == 2)
        bp_forward_kernel<2><<<grid, 64 * BP_WAVES, 0, stream>>>(proj, frames, g, chunk, scale, clamp, vol);
    else
        bp_forward_kernel<1><<<grid, 64 * BP_WAVES, 0, stream>>>(proj, frames, g, chunk, scale, clamp, vol);
    CBCT_LAUNCHED("backprojection kernel launch");
    return CBCT_OK;
}

int launch_bp_forward_fdk(const float* proj, const ViewFrame* frames, int64_t A, int64_t U, int64_t V,
                          int64_t D, int64_t H, int64_t W, float pix, float vs, float ox, float oy,
                          float oz, int chunk, float scale, float* vol, hipStream_t stream) {
    const BpGeom g = make_bp_geom(A, U, V, D, H, W, pix, vs, ox, oy, oz);
    return launch_bp_forward(proj, frames, g, chunk, scale, 1, vol, stream);
}

}  // namespace cbct

using namespace cbct;

extern "C" size_t cbct_bp_workspace_bytes(int64_t A, int64_t D, int64_t H, int64_t W) {
    (void)D; (void)H; (void)W;
    if (A < 1) return 0;
    return (size_t)A * sizeof(ViewFrame);
}

// frames | per-workgroup geometry partials [A][blocks][8] | (deterministic) int64 accumulator
// [A,U,V] | max bits + per-view exponents
extern "C" size_t cbct_bp_backward_workspace_bytes_ex(int64_t A, int64_t U, int64_t V, int64_t D, int64_t H,
                                                      int64_t W, const cbct_fp_options* options) {
    if (A < 1 || U < 1 || V < 1 || D < 1 || H < 1 || W < 1) return 0;
    int det = 0;
    if (call_deterministic(options, &det)) return 0;
    const dim3 grid = bp_bwd_grid(D, H, W);
    size_t n = align256((size_t)A * sizeof(ViewFrame)) + align256((size_t)A * grid.x * grid.y * 8 * sizeof(float));
    if (det) n += align256((size_t)A * U * V * sizeof(unsigned long long)) + align256(sizeof(unsigned) * (1 + A));
    return n;
}

extern "C" size_t cbct_bp_backward_workspace_bytes(int64_t A, int64_t U, int64_t V, int64_t D, int64_t H,
                                                   int64_t W) {
    return cbct_bp_backward_workspace_bytes_ex(A, U, V, D, H, W, nullptr);
}

extern "C" int cbct_bp_forward(const float* projections, int64_t A, int64_t U, int64_t V,
                               const float* angles, const float* sid, const float* sod,
                               const float* det_u0, const float* det_v0,
                               const float* src_x, const float* src_y, const float* src_z,
                               float pixel_size, int64_t D, int64_t H, int64_t W, float voxel_size,
                               float off_x, float off_y, float off_z,
                               float* volume, void* workspace, size_t workspace_bytes, void* stream_) {
    int st;
    if ((st = check_bp_common(A, U, V, D, H, W, pixel_size, voxel_size, off_x, off_y, off_z))) return st;
    const float* gg[8] = {angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z};
    if ((st = check_geometry(gg, A))) return st;
    CBCT_CHECK_ARG(projections != nullptr && volume != nullptr, "projections and volume must be non-null");
    CBCT_CHECK_ARG(workspace != nullptr && workspace_bytes >= cbct_bp_workspace_bytes(A, D, H, W),
                   "workspace too small for cbct_bp_forward");
    hipStream_t stream = (hipStream_t)stream_;
    ViewFrame* frames = (ViewFrame*)workspace;
    const GeomPtrs gp{angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z};
    if ((st = launch_frames(gp, A, frames, stream))) return st;
    const BpGeom g = make_bp_geom(A, U, V, D, H, W, pixel_size, voxel_size, off_x, off_y, off_z);
    return launch_bp_forward(projections, frames, g, (int)A, 1.0f, 0, volume, stream);
}

extern "C" int cbct_bp_backward_ex(const float* grad_volume, const float* projections,
                                   int64_t A, int64_t U, int64_t V,
                                   const float* angles, const float* sid, const float* sod,
                                   const float* det_u0, const float* det_v0,
                                   const float* src_x, const float* src_y, const float* src_z,
                                   float pixel_size, int64_t D, int64_t H, int64_t W, float voxel_size,
                                   float off_x, float off_y, float off_z,
                                   float* grad_projections, float* grad_geometry, const cbct_fp_options* options,
                                   void* workspace, size_t workspace_bytes, void* stream_) {
    int st;
    if ((st = check_bp_common(A, U, V, D, H, W, pixel_size, voxel_size, off_x, off_y, off_z))) return st;
    const float* gg[8] = {angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z};
    if ((st = check_geometry(gg, A))) return st;
    CBCT_CHECK_ARG(grad_volume != nullptr && projections != nullptr && grad_projections != nullptr &&
                   grad_geometry != nullptr, "gradient and projection pointers must be non-null");
    int det_flag = 0;
    if ((st = call_deterministic(options, &det_flag))) return st;
    CBCT_CHECK_ARG(workspace != nullptr &&
                   workspace_bytes >= cbct_bp_backward_workspace_bytes_ex(A, U, V, D, H, W, options),
                   "workspace too small for cbct_bp_backward (see cbct_bp_backward_workspace_bytes)");
    const bool det = det_flag != 0;
    hipStream_t stream = (hipStream_t)stream_;
    const dim3 grid = bp_bwd_grid(D, H, W);
    const int64_t nblk = (int64_t)grid.x * grid.y;
    ViewFrame* frames = (ViewFrame*)workspace;
    float* partials = (float*)((char*)workspace + align256((size_t)A * sizeof(ViewFrame)));
    const GeomPtrs gp{angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z};
    if ((st = launch_frames(gp, A, frames, stream))) return st;
    const BpGeom g = make_bp_geom(A, U, V, D, H, W, pixel_size, voxel_size, off_x, off_y, off_z);
    unsigned long long* gacc = nullptr;
    int* gexp = nullptr;
    const int64_t nproj = A * U * V;
    if (det) {
        gacc = (unsigned long long*)((char*)partials + align256((size_t)A * nblk * 8 * sizeof(float)));
        unsigned* maxbits = (unsigned*)((char*)gacc + align256((size_t)nproj * sizeof(unsigned long long)));
        gexp = (int*)(maxbits + 1);
        if ((st = launch_zero_u64(gacc, nproj, stream))) return st;
        if ((st = launch_fill((float*)maxbits, 1, 0.0f, stream))) return st;
        if ((st = launch_absmax(grad_volume, D * H * W, maxbits, stream))) return st;
        bp_det_exponent_kernel<<<(unsigned)A, 256, 0, stream>>>(frames, g, maxbits, gexp);
        CBCT_LAUNCHED("deterministic backward exponent kernel");
    } else if ((st = launch_fill(grad_projections, nproj, 0.0f, stream))) {
        return st;
    }
    if (bp_bwd_k(D) == 2) {
        if (det)
            bp_backward_kernel<2, true><<<grid, 64 * BP_WAVES, 0, stream>>>(grad_volume, projections, frames, g,
                                                                           grad_projections, partials, gacc, gexp);
        else
            bp_backward_kernel<2, false><<<grid, 64 * BP_WAVES, 0, stream>>>(grad_volume, projections, frames, g,
                                                                            grad_projections, partials, gacc, gexp);
    } else {
        if (det)
            bp_backward_kernel<1, true><<<grid, 64 * BP_WAVES, 0, stream>>>(grad_volume, projections, frames, g,
                                                                           grad_projections, partials, gacc, gexp);
        else
            bp_backward_kernel<1, false><<<grid, 64 * BP_WAVES, 0, stream>>>(grad_volume, projections, frames, g,
                                                                            grad_projections, partials, gacc, gexp);
    }
    CBCT_LAUNCHED("backprojection backward kernel launch");
    if (det) {
        bp_det_convert_kernel<<<(unsigned)ceil_div(nproj, (int64_t)256), 256, 0, stream>>>(gacc, nproj, U * V, gexp,
                                                                                         grad_projections);
        CBCT_LAUNCHED("deterministic projection gradient conversion");
    }
    return launch_reduce_geometry(partials, A, nblk, grad_geometry, stream);
}

extern "C" int cbct_bp_backward(const float* grad_volume, const float* projections,
                                int64_t A, int64_t U, int64_t V,
                                const float* angles, const float* sid, const float* sod,
                                const float* det_u0, const float* det_v0,
                                const float* src_x, const float* src_y, const float* src_z,
                                float pixel_size, int64_t D, int64_t H, int64_t W, float voxel_size,
                                float off_x, float off_y, float off_z,
                                float* grad_projections, float* grad_geometry,
                                void* workspace, size_t workspace_bytes, void* stream_) {
    return cbct_bp_backward_ex(grad_volume, projections, A, U, V, angles, sid, sod, det_u0, det_v0, src_x, src_y,
                               src_z, pixel_size, D, H, W, voxel_size, off_x, off_y, off_z, grad_projections,
                               grad_geometry, nullptr, workspace, workspace_bytes, stream_);
}

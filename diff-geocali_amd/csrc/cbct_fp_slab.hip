// Slab-staged forward projector and adjoint for gfx950.
//
// A workgroup owns a 16x16 tile of detector pixels of one view (one ray per
// lane).  The tile's rays are marched together through the volume in slabs of
// T voxel planes along the tile's principal axis (the axis the central ray
// moves fastest along).  For each slab the workgroup
//   1. bounds the slab's part of the tile's ray bundle: the convex hull of the
//      four corner rays between the slab's two planes, widened by the trilinear
//      support and a rounding margin;
//   2. stages that box of the volume in LDS (and, for the adjoint, zeroes a
//      second LDS box that accumulates the volume gradient);
//   3. advances every ray through the samples whose principal voxel
//      coordinate lies in the slab, reading the eight taps from LDS (and adding
//      the eight weighted contributions to the LDS gradient box);
//   4. flushes the gradient box to HBM with one atomic add per non-zero voxel.
// The arithmetic of every sample is the same float code as the direct kernels
// (same t schedule, same cells, same interpolation), so projections and geometry
// gradients are bit-identical to them; only the order of the volume-gradient
// atomics differs.  A ray that cannot show that its samples of a slab have their
// taps in the staged box (never in a regular cone-beam geometry) and tiles whose
// rays do not share one principal direction take the direct path from global
// memory, so the staging only ever affects speed.
#include "cbct_fp_slab.h"

#include <algorithm>
#include <type_traits>

namespace cbct {

__device__ __forceinline__ void pixel_dir(const ViewFrame& f, float u, float v, float pix, float r[3]) {
    const float up = (u - f.u0) * pix, vp = (v - f.v0) * pix;
    r[0] = fmaf(up, f.c, -f.ns) - f.ox;
    r[1] = fmaf(up, f.s, f.nc) - f.oy;
    r[2] = vp - f.oz;
}

// One 16x16 tile of rays of one problem.
// MODE 0: forward; 1: adjoint, geometry gradients only; 2: adjoint with the volume gradient;
// 3: calibration step -- forward, squared residual against the measured projections (P.grad) and
//    the geometry gradient of the loss, in one march (the geometry sums are linear in the upstream
//    gradient, so they are accumulated with unit weight and scaled by the residual at the end).
// 4: Gauss-Newton step -- as 3, but each ray's eight geometry derivatives j of its projection and
//    its residual r give the tile's normal equations: the 36 entries of sum j j^T (upper triangle,
//    row-major), the 8 of sum j r and the loss (45 partials per tile).
// GEO (MODE 2 only): with false, the geometry gradients are not wanted -- the adjoint then needs
// no volume taps at all (the scatter depends only on the sample positions and the upstream
// weights): no staged volume box, no interpolation, no geometry sums, one LDS box per workgroup.
// BIG: the volume has 2^30 or more voxels, so voxel offsets are formed in 64 bits (the
// 32-bit forms of the other variants are cheaper in the staging and flush loops).
template <int MODE, bool DIV, bool ZOFF = false, int CAPC = 0, bool GEO = true, bool BIG = false>
__device__ __forceinline__ void slab_tile(const SlabArgs& P, int tile) {
    using VoxelIndex = typename std::conditional<BIG, int64_t, int>::type;
    static_assert(GEO || MODE == 2, "only the volume-gradient adjoint can skip the geometry");
    constexpr bool TAPS = GEO;        // the march reads volume taps
    extern __shared__ float lds[];
    float* vbox = lds;
    // CAPC > 0: the LDS capacity as a compile-time constant (== P.cap), so the gradient box's
    // addresses are the tap addresses plus an immediate offset
    float* gbox = TAPS ? lds + (CAPC > 0 ? CAPC : P.cap) : lds;
    const Box& b = P.b;

    // ---- tile and ray of this lane
    const int tv = tile % P.tiles_v;
    const int rest = tile / P.tiles_v;
    const int tu = rest % P.tiles_u;
    const int a = rest / P.tiles_u;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int du, dv;
    if (P.layout == 0) { dv = ((wave & 1) << 3) + (lane & 7); du = ((wave >> 1) << 3) + (lane >> 3); }
    else { dv = ((lane & 7) << 1) + (wave & 1); du = ((lane >> 3) << 1) + (wave >> 1); }
    const int u = tu * SL_TILE + du, v = tv * SL_TILE + dv;
    const bool inr = u < P.U && v < P.V;
    const ViewFrame f = P.frames[a];

    Ray ray;
    float tmin = 0.f, tmax = -1.f, t = 0.f;
    int entry = -1;
    bool alive = false, valid = false, hit = false;
    float go = 0.f;
    if (inr) {
        if (MODE != 0) go = P.grad[((int64_t)a * P.U + u) * P.V + v];
        if ((MODE == 0 || MODE >= 3 || go != 0.f) && make_ray(f, u, v, P.pix, ray)) {
            valid = true;
            if (intersect(b, f, ray, tmin, tmax, entry)) { hit = true; alive = true; t = tmin; }
            else tmax = -1.f;        // keep t <= tmax equivalent to alive (t = 0) for a miss
        }
    }
    const float gs = MODE >= 3 ? 1.0f : go * P.step;
    float integral = 0.f;
    // Fixed-point scale of the LDS gradient box (MODE 2): a power of two 2^e chosen so that the
    // largest total a voxel can receive from this tile within one slab stays below 2^30.  That
    // total is at most max|gs| * n_rays * n_smp, where n_smp bounds the samples of one ray inside
    // a voxel's 2x2x2 support (floor(2 sqrt3 vs / step) + 1) and n_rays bounds the tile's rays
    // through that support (pixels within its projected radius, or all 256).
    float gsq = 0.f, hq = 0.f, inv_scale = 1.f;
    int e_tile = 0;                                   // the tile's fixed-point exponent
    const int E = (MODE == 2 && P.gacc) ? *P.gexp : 0;  // the global exponent (deterministic mode)
    const double scaleE = ldexp(1.0, E);
    __shared__ float red_max[SL_THREADS / 64];
    unsigned* gboxu = reinterpret_cast<unsigned*>(gbox);
    if (MODE == 2) {
        float m = fabsf(gs);
        for (int s = 32; s >= 1; s >>= 1) m = fmaxf(m, __shfl_xor(m, s, 64));
        if ((threadIdx.x & 63) == 0) red_max[threadIdx.x >> 6] = m;
        __syncthreads();
        const float amax = fmaxf(fmaxf(red_max[0], red_max[1]), fmaxf(red_max[2], red_max[3]));
        // distance from the source to the volume box and the widest ray angle of the tile
        const float ddx = fmaxf(fmaxf(b.lox - f.ox, f.ox - b.hix), 0.f);
        const float ddy = fmaxf(fmaxf(b.loy - f.oy, f.oy - b.hiy), 0.f);
        const float ddz = fmaxf(fmaxf(b.loz - f.oz, f.oz - b.hiz), 0.f);
        const float dmin = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
        float cmin = 1.f;
        {
            const int ul = tu * SL_TILE, vl = tv * SL_TILE;
            const float us[2] = {(float)ul, (float)min(ul + SL_TILE - 1, P.U - 1)};
            const float vv[2] = {(float)vl, (float)min(vl + SL_TILE - 1, P.V - 1)};
            for (int i = 0; i < 4; ++i) {
                float r[3];
                pixel_dir(f, us[i & 1], vv[i >> 1], P.pix, r);
                const float rn = (-f.s * r[0] + f.c * r[1]);     // component along the detector normal
                const float len = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
                cmin = fminf(cmin, fabsf(rn) / len);
            }
        }
        float n_rays = (float)SL_THREADS;
        if (dmin > 0.f && cmin > 0.2f) {
            const float rpx = 1.05f * 1.7320508f * P.vs * f.sid / (dmin * P.pix * cmin * cmin);
            n_rays = fminf(n_rays, (2.f * rpx + 2.f) * (2.f * rpx + 2.f));
        }
        const float n_smp = floorf(2.f * 1.7320508f * P.vs / P.step) + 1.f;
        const float bound = amax * n_rays * n_smp;
        int e = 0;
        if (bound > 0.f) {
            int k;
            frexpf(1073741824.f / bound, &k);     // 2^30 / bound = m * 2^k, m in [0.5, 1)
            e = max(-100, min(100, k - 1));
        }
        gsq = gs * ldexpf(1.f, e);
        hq = copysignf(0.5f, gsq);
        inv_scale = ldexpf(1.f, -e);
        e_tile = e;
    }
    float gox = 0.f, goy = 0.f, goz = 0.f, gdx = 0.f, gdy = 0.f, gdz = 0.f, gtmin = 0.f;

    // The reciprocal voxel size lives in a VGPR so that fmaf(p - off, inv_vs, half) takes one scalar
    // operand (half) and needs no per-sample copy of the other.
    float inv_vs_v = P.inv_vs;
    asm volatile("" : "+v"(inv_vs_v));
    auto voxel_coords_at = [&](float tt, float& x, float& y, float& z) {
        const float px = fmaf(tt, ray.dx, f.ox), py = fmaf(tt, ray.dy, f.oy), pz = fmaf(tt, ray.dz, f.oz);
        // ZOFF: all volume offsets are zero, and p - 0 == p exactly, so the subtraction folds away
        const float ox_ = ZOFF ? 0.f : b.offx, oy_ = ZOFF ? 0.f : b.offy, oz_ = ZOFF ? 0.f : b.offz;
        if (MODE == 0 || MODE >= 3) {      // the forward's voxel map (division unless vs is a power of two)
            x = to_voxel<DIV>(px, ox_, P.vs, inv_vs_v, b.hx);
            y = to_voxel<DIV>(py, oy_, P.vs, inv_vs_v, b.hy);
            z = to_voxel<DIV>(pz, oz_, P.vs, inv_vs_v, b.hz);
        } else {
            x = fmaf(px - ox_, inv_vs_v, b.hx);
            y = fmaf(py - oy_, inv_vs_v, b.hy);
            z = fmaf(pz - oz_, inv_vs_v, b.hz);
        }
    };
    auto voxel_coords = [&](float& x, float& y, float& z) { voxel_coords_at(t, x, y, z); };
    // Everything a sample does once its eight taps are known.  Lanes with go == false (no sample in
    // the current slab) leave every sum unchanged.
    auto use_taps = [&](const Taps& tp, float fx, float fy, float fz, bool go) {
        if (!TAPS) return;
        float val, sgx, sgy, sgz;
        tri_eval(tp, fx, fy, fz, val, sgx, sgy, sgz);
        if (MODE == 0 || MODE >= 3) integral += go ? val : 0.f;
        if (MODE != 0) {
            // unscaled sums of the voxel-unit spatial gradient and of its product with t; the
            // ray's constant factors (upstream weight, step, 1/voxel size) are applied once after
            // the march, and the t_min sum is their dot product with the direction
            sgx = go ? sgx : 0.f; sgy = go ? sgy : 0.f; sgz = go ? sgz : 0.f;
            gox += sgx; goy += sgy; goz += sgz;
            gdx = fmaf(sgx, t, gdx); gdy = fmaf(sgy, t, gdy); gdz = fmaf(sgz, t, gdz);
        }
    };
    const int plane = b.H * b.W;
    // The taps of a sample from global memory with the reference's clamp-to-edge corners (and, for
    // the adjoint, its eight volume-gradient contributions).
    auto global_taps = [&](float x, float y, float z) {
        const Cell c = cell_of(b, x, y, z);
        const VoxelIndex i000 = (VoxelIndex)(c.z0 * b.H + c.y0) * b.W + c.x0;   // z0*H + y0 < D*H < 2^31
        const int ox1 = c.sx, oy1 = c.sy * b.W, oz1 = c.sz * plane;
        Taps tp{};
        if (TAPS) tp = load_taps_strided(P.vol, i000, ox1, oy1, oz1);
        if (MODE == 2 && P.gacc) {
            const Weights w = tri_weights(c.fx, c.fy, c.fz);
            unsigned long long* q = P.gacc + i000;
            auto add = [&](unsigned long long* dst, float v) {
                const long long iv = (long long)rint((double)v * scaleE);
                if (iv != 0) atomicAdd(dst, (unsigned long long)iv);
            };
            add(q, gs * w.w000);
            add(q + ox1, gs * w.w100);
            add(q + oy1, gs * w.w010);
            add(q + oy1 + ox1, gs * w.w110);
            add(q + oz1, gs * w.w001);
            add(q + oz1 + ox1, gs * w.w101);
            add(q + oz1 + oy1, gs * w.w011);
            add(q + oz1 + oy1 + ox1, gs * w.w111);
        } else if (MODE == 2) {
            const Weights w = tri_weights(c.fx, c.fy, c.fz);
            float* q = P.gvol + i000;
            atomicAdd(q, gs * w.w000);
            atomicAdd(q + ox1, gs * w.w100);
            atomicAdd(q + oy1, gs * w.w010);
            atomicAdd(q + oy1 + ox1, gs * w.w110);
            atomicAdd(q + oz1, gs * w.w001);
            atomicAdd(q + oz1 + ox1, gs * w.w101);
            atomicAdd(q + oz1 + oy1, gs * w.w011);
            atomicAdd(q + oz1 + oy1 + ox1, gs * w.w111);
        }
        return tp;
    };
    auto global_step = [&]() {
        float x, y, z;
        voxel_coords(x, y, z);
        const Taps tp = global_taps(x, y, z);
        use_taps(tp, frac(x), frac(y), frac(z), true);
    };

    // (axis k of a 3-vector by selects: a runtime array index would put the array in scratch)
    auto pick = [](const float* v3, int k) { return k == 0 ? v3[0] : (k == 1 ? v3[1] : v3[2]); };
    auto pick_i = [](const int* v3, int k) { return k == 0 ? v3[0] : (k == 1 ? v3[1] : v3[2]); };
    // ---- principal axis of the tile (uniform over the workgroup)
    const int u_lo = tu * SL_TILE, v_lo = tv * SL_TILE;
    const int u_hi = min(u_lo + SL_TILE - 1, P.U - 1), v_hi = min(v_lo + SL_TILE - 1, P.V - 1);
    float cr[3];
    pixel_dir(f, 0.5f * (float)(u_lo + u_hi), 0.5f * (float)(v_lo + v_hi), P.pix, cr);
    int p = 0;
    if (fabsf(cr[1]) > fabsf(cr[0])) p = 1;
    if (fabsf(cr[2]) > fabsf(pick(cr, p))) p = 2;
    // every lane computed the same axis and direction; make that explicit (scalar registers,
    // uniform branches in the march loop)
    p = __builtin_amdgcn_readfirstlane(p);
    const int fwd = __builtin_amdgcn_readfirstlane(pick(cr, p) > 0.f ? 1 : 0);
    const float dirf = fwd ? 1.f : -1.f;
    float corner[4][3];
    pixel_dir(f, (float)u_lo, (float)v_lo, P.pix, corner[0]);
    pixel_dir(f, (float)u_hi, (float)v_lo, P.pix, corner[1]);
    pixel_dir(f, (float)u_lo, (float)v_hi, P.pix, corner[2]);
    pixel_dir(f, (float)u_hi, (float)v_hi, P.pix, corner[3]);
    bool slab_ok = true;
    for (int j = 0; j < 4; ++j) {
        const float n2 = corner[j][0] * corner[j][0] + corner[j][1] * corner[j][1] + corner[j][2] * corner[j][2];
        const float cp = pick(corner[j], p) * dirf;
        slab_ok = slab_ok && cp > 0.f && cp * cp >= 0.0625f * n2;
    }
    const int Nax[3] = {b.W, b.H, b.D};
    const float offs[3] = {b.offx, b.offy, b.offz};
    const float halfs[3] = {b.hx, b.hy, b.hz};
    const float src[3] = {f.ox, f.oy, f.oz};
    float xs[3];
    for (int i = 0; i < 3; ++i) xs[i] = (src[i] - offs[i]) * P.inv_vs + halfs[i];
    const int o1 = p == 0 ? 1 : 0, o2 = p == 2 ? 1 : 2;

    const float dpr = p == 0 ? ray.dx : (p == 1 ? ray.dy : ray.dz);
    const bool in_slabs = alive && slab_ok && dpr * dirf > 0.f;
    // rays that cannot follow the slab order are integrated from global memory first
    if (alive && !in_slabs) {
        while (alive) {
            global_step();
            t += P.step;
            alive = t <= tmax;
        }
    }

    // Box of the tile's samples whose principal floor coordinate lies in [s_lo, s_hi], in raw
    // (unclamped) cell coordinates: the hull of the corner rays between the planes q = s_lo and
    // q = s_hi + 1, widened by the +1 taps and a small rounding margin, limited to [-1, N].  The staged
    // copy holds vol[clamp(.)] at every raw coordinate, so taps read from it need no clamping.
    // Returns false when no cell fits in it.
    // Lateral slopes of the four corner rays (voxel units per principal voxel), once per tile:
    // between the planes q = s_lo and q = s_hi + 1 a corner ray's lateral coordinates are
    // xs + (q - xs_p) * slope.  Every corner ray moves in the march direction (slab_ok), so the
    // planes lie in front of the source exactly when (q - xs_p) has the direction's sign.
    float sl1[4], sl2[4];
    for (int j = 0; j < 4; ++j) {
        const float cpj = pick(corner[j], p);
        sl1[j] = pick(corner[j], o1) / cpj;
        sl2[j] = pick(corner[j], o2) / cpj;
    }
    const float xsp = pick(xs, p), xs1 = pick(xs, o1), xs2 = pick(xs, o2);
    const int Nx1 = pick_i(Nax, o1), Nx2 = pick_i(Nax, o2);
    const int Np = pick_i(Nax, p);
    auto slab_box = [&](int s_lo, int s_hi, int lo[3], int hi[3]) {
        const float q0 = (float)s_lo - xsp, q1 = (float)(s_hi + 1) - xsp;
        const bool front = q0 * dirf > 0.f && q1 * dirf > 0.f;
        float mn1 = FLT_MAX, mx1 = -FLT_MAX, mn2 = FLT_MAX, mx2 = -FLT_MAX;
        for (int j = 0; j < 4; ++j) {
            const float a1 = fmaf(q0, sl1[j], xs1), b1 = fmaf(q1, sl1[j], xs1);
            const float a2 = fmaf(q0, sl2[j], xs2), b2 = fmaf(q1, sl2[j], xs2);
            mn1 = fminf(mn1, fminf(a1, b1)); mx1 = fmaxf(mx1, fmaxf(a1, b1));
            mn2 = fminf(mn2, fminf(a2, b2)); mx2 = fmaxf(mx2, fmaxf(a2, b2));
        }
        const int lp = max(-1, s_lo), hp = min(Np, s_hi + 1);
        // rounding margin: 0.01 voxel, far above the hull's FP32 error (~1e-4 voxel); a sample
        // that still fell outside would only send its lane to the global path (in-box proof)
        const int l1 = (int)fminf(fmaxf(floorf(mn1 - 0.01f), -1.f), (float)Nx1);
        const int h1 = (int)fmaxf(fminf(floorf(mx1 + 0.01f) + 1.f, (float)Nx1), -1.f);
        const int l2 = (int)fminf(fmaxf(floorf(mn2 - 0.01f), -1.f), (float)Nx2);
        const int h2 = (int)fmaxf(fminf(floorf(mx2 + 0.01f) + 1.f, (float)Nx2), -1.f);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = k == p ? lp : (k == o1 ? l1 : l2);
            hi[k] = k == p ? hp : (k == o1 ? h1 : h2);
        }
        if (MODE == 2 && !P.gacc) {
            // odd row and plane pitches for the adjoint's boxes: neighbouring lanes' eight
            // atomics then spread over more LDS banks (171 vs 176 ms at 256^3 x 720; not in
            // deterministic mode, whose int64 flush of the larger boxes costs more: 207 vs 199 ms)
            hi[0] += (hi[0] - lo[0] + 1) & 1 ? 0 : 1;
            hi[1] += (hi[1] - lo[1] + 1) & 1 ? 0 : 1;
        }
        return front && lo[0] < hi[0] && lo[1] < hi[1] && lo[2] < hi[2];
    };

    // the step in a VGPR: the march's t update is then one select and one add (not in the
    // adjoint, which has no register to spare for it)
    float step_v = P.step;
    if (MODE != 2) asm volatile("" : "+v"(step_v));

    // The samples of one slab: those whose principal floor coordinate lies in [s_lo, s_hi], with
    // the taps read from the staged box.  PAX / FWD (the tile's principal axis and direction) are
    // compile-time constants, so the slab-exit test is one compare.  The loop runs while any lane
    // of the wave has a sample in the slab; a lane without one rides along with its updates masked
    // (go == false), so the loop has no divergent exits.  A lane that cannot show that all its
    // samples of the slab have their taps in the box (not seen in regular cone-beam geometry;
    // every lane when the slab is not staged) reads them from global memory.
    auto march = [&](auto pax_c, auto fwd_c, int s_lo, int s_hi, const int* lo, int B0, int B1, int B2,
                     bool staged) {
        constexpr int PAX = decltype(pax_c)::value;
        constexpr bool FWD = decltype(fwd_c)::value;
        const unsigned bx1 = staged ? (unsigned)(B0 - 1) : 0u, by1 = staged ? (unsigned)(B1 - 1) : 0u;
        const unsigned bz1 = staged ? (unsigned)(B2 - 1) : 0u;
        const int PL = B0 * B1;
        // byte offset of the box's first cell in cell coordinates
        const int base4 = 4 * (lo[2] * PL + lo[1] * B0 + lo[0]);
        const int B04 = 4 * B0, PL4 = 4 * PL;
        // Per lane, once per slab: do the taps of every sample of this slab lie in the box?  The
        // lane's samples here lie between its current t and t_end = min(t_x, tmax), where t_x is a
        // t at which the principal coordinate has left the slab (solved for, moved on by 0.01
        // voxel, and checked).  Each voxel coordinate is a monotone function of t (fmaf and the
        // voxel map are correctly rounded, with fixed signs), so if the cells at both ends have
        // their taps in the box, so do all samples in between, and no sample needs its own test.
        // A lane that fails takes the global path for this slab.
        bool lane_safe = false;
        if (staged && t <= tmax) {
            float x, y, z;
            voxel_coords(x, y, z);
            bool ok = (unsigned)(floor_int(x) - lo[0]) < bx1 && (unsigned)(floor_int(y) - lo[1]) < by1 &&
                      (unsigned)(floor_int(z) - lo[2]) < bz1;
            const float dp = PAX == 0 ? ray.dx : (PAX == 1 ? ray.dy : ray.dz);
            const float op = PAX == 0 ? f.ox : (PAX == 1 ? f.oy : f.oz);
            const float offp = PAX == 0 ? b.offx : (PAX == 1 ? b.offy : b.offz);
            const float hp = PAX == 0 ? b.hx : (PAX == 1 ? b.hy : b.hz);
            const float target = (float)(FWD ? s_hi + 1 : s_lo);
            const float tx = (fmaf(target - hp, P.vs, offp) - op) / dp + 0.01f * P.vs / fabsf(dp);
            float te = tmax;
            if (tx < tmax) {
                te = tx;
                voxel_coords_at(tx, x, y, z);
                const float cp = PAX == 0 ? x : (PAX == 1 ? y : z);
                ok = ok && (FWD ? cp >= target : cp < target);
            }
            voxel_coords_at(te, x, y, z);
            ok = ok && (PAX == 0 || (unsigned)(floor_int(x) - lo[0]) < bx1) &&
                 (PAX == 1 || (unsigned)(floor_int(y) - lo[1]) < by1) &&
                 (PAX == 2 || (unsigned)(floor_int(z) - lo[2]) < bz1);
            lane_safe = ok;
        }
        for (;;) {
            float x, y, z;
            voxel_coords(x, y, z);
            const int xr = floor_int(x), yr = floor_int(y), zr = floor_int(z);
            const int fp = PAX == 0 ? xr : (PAX == 1 ? yr : zr);
            // A lane is alive while t <= tmax (never, for a ray that misses the volume).  Lane
            // predicates are ands of compares; the loop exit is a ballot of the same compares
            // (scalar registers only).
            const bool live = t <= tmax;
            const bool ins = FWD ? fp <= s_hi : fp >= s_lo;
            const uint64_t m_go = __builtin_amdgcn_ballot_w64(live) & __builtin_amdgcn_ballot_w64(ins);
            if (m_go == 0ull) break;
            const bool go = live && ins;
            const bool gl = go && lane_safe;
            const float fx = frac(x), fy = frac(y), fz = frac(z);
            if (gl) {
                // (gl: the slab is staged and the lane's samples in it are in the box; the other
                // lanes skip the block.)  Byte offset of the cell in the box: 24-bit products (cell
                // coordinates and box strides are small), full rate.
                const unsigned idx4 = (unsigned)(__mul24(zr, PL4) + __mul24(yr, B04) + 4 * xr - base4);
                if (TAPS) {
                    const Taps tp = load_taps_strided(reinterpret_cast<const float*>(
                                                          reinterpret_cast<const char*>(vbox) + idx4), 0, 1, B0, PL);
                    use_taps(tp, fx, fy, fz, true);
                }
                if (MODE == 2) {
                    // the scaled trilinear weights as a z, then y split of gsq, then x weights
                    // plus half a unit of gsq's sign (every weight has it), so that truncation
                    // (one conversion) rounds to nearest, halves away from zero
                    const float g1 = gsq * fz, g0 = gsq - g1;
                    const float g01 = g0 * fy, g00 = g0 - g01, g11 = g1 * fy, g10 = g1 - g11;
                    const float mx = 1.0f - fx;
                    unsigned* qi = reinterpret_cast<unsigned*>(reinterpret_cast<char*>(gboxu) + idx4);
                    atomicAdd(qi, (unsigned)(int)fmaf(g00, mx, hq));
                    atomicAdd(qi + 1, (unsigned)(int)fmaf(g00, fx, hq));
                    atomicAdd(qi + B0, (unsigned)(int)fmaf(g01, mx, hq));
                    atomicAdd(qi + B0 + 1, (unsigned)(int)fmaf(g01, fx, hq));
                    atomicAdd(qi + PL, (unsigned)(int)fmaf(g10, mx, hq));
                    atomicAdd(qi + PL + 1, (unsigned)(int)fmaf(g10, fx, hq));
                    atomicAdd(qi + PL + B0, (unsigned)(int)fmaf(g11, mx, hq));
                    atomicAdd(qi + PL + B0 + 1, (unsigned)(int)fmaf(g11, fx, hq));
                }
            }
            if (go && !gl) {      // a divergent branch that no lane takes costs a skip
                const Taps tg = global_taps(x, y, z);
                use_taps(tg, fx, fy, fz, true);
            }
            if (MODE != 2) t += go ? step_v : 0.f;
            else t = go ? t + P.step : t;
        }
        alive = t <= tmax;
    };

    // (a tile none of whose rays meets the volume in slab order has nothing to stage)
    const bool any_ray = __syncthreads_or(in_slabs ? 1 : 0) != 0;
    if (slab_ok && any_ray) {
        // Slabs cover the principal floor coordinates -1 .. Np-1 in the rays' direction; a slab is
        // halved (to an even thickness) until its box fits the LDS capacity (uniform over the workgroup).
        // The gathering kernels let wave 0 bound each slab and pass the result to the other waves
        // through the last SLAB_MSG words of the box's LDS, which no box uses (two buffers,
        // alternating: a buffer is rewritten only after the barrier that every wave passes after
        // reading it; a static LDS array would cost the forward a workgroup per CU).  That barrier
        // also orders the previous slab's march before this slab's staging.  The adjoint, whose
        // workgroups already wait for each flush, bounds the slab in every wave instead (faster
        // there: 171 vs 176 ms, while the forward gains 74.3 -> 73.2 ms).
        constexpr bool BCAST = MODE != 2;
        constexpr int SLAB_MSG = BCAST ? 20 : 0;
        const int box_cap = (CAPC > 0 ? CAPC : P.cap) - SLAB_MSG;
        int* const slab_sh = reinterpret_cast<int*>(vbox + box_cap);
        int parity = 0;
        int next = dirf > 0.f ? -1 : Np - 1;
        // the slab starting at next, halved (to an even thickness) until its box fits
        auto bound_slab = [&](int& thick, int& s_lo, int& s_hi, int* lo, int* hi) {
            thick = P.T;
            bool nonempty;
            for (;;) {
                if (dirf > 0.f) { s_lo = next; s_hi = next + thick - 1; }
                else { s_hi = next; s_lo = next - thick + 1; }
                nonempty = slab_box(s_lo, s_hi, lo, hi);
                // identical in every lane: keep the box in scalar registers
                nonempty = __builtin_amdgcn_readfirstlane(nonempty ? 1 : 0) != 0;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    lo[i] = __builtin_amdgcn_readfirstlane(lo[i]);
                    hi[i] = __builtin_amdgcn_readfirstlane(hi[i]);
                }
                if (!nonempty || thick == 1 ||
                    (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1) <= box_cap)
                    break;
                // halve, to an even thickness: the box's extent along the principal axis (the
                // row pitch when that is x) is then odd, which keeps neighbouring lanes' taps
                // on different LDS banks
                thick = max(1, ((thick + 1) >> 1) & ~1);
            }
            return nonempty;
        };
        while (dirf > 0.f ? next <= Np - 1 : next >= -1) {
            int thick, s_lo, s_hi;
            int lo[3], hi[3];
            bool nonempty;
            if constexpr (BCAST) {
                if (wave == 0) {
                    nonempty = bound_slab(thick, s_lo, s_hi, lo, hi);
                    if (lane == 0) {
                        int* o = slab_sh + 10 * parity;
                        o[0] = thick; o[1] = s_lo; o[2] = s_hi; o[3] = nonempty ? 1 : 0;
                        o[4] = lo[0]; o[5] = lo[1]; o[6] = lo[2]; o[7] = hi[0]; o[8] = hi[1]; o[9] = hi[2];
                    }
                }
                __syncthreads();
                const int* o = slab_sh + 10 * parity;
                parity ^= 1;
                thick = __builtin_amdgcn_readfirstlane(o[0]);
                s_lo = __builtin_amdgcn_readfirstlane(o[1]);
                s_hi = __builtin_amdgcn_readfirstlane(o[2]);
                nonempty = __builtin_amdgcn_readfirstlane(o[3]) != 0;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    lo[i] = __builtin_amdgcn_readfirstlane(o[4 + i]);
                    hi[i] = __builtin_amdgcn_readfirstlane(o[7 + i]);
                }
            } else {
                nonempty = bound_slab(thick, s_lo, s_hi, lo, hi);
            }
            next += dirf > 0.f ? thick : -thick;
            const bool empty = !nonempty;
            const int B0 = hi[0] - lo[0] + 1, B1 = hi[1] - lo[1] + 1, B2 = hi[2] - lo[2] + 1;
            const bool staged = !empty && B0 * B1 * B2 <= box_cap;
            if (!staged && empty) continue;   // no sample of the tile can lie in this slab
            const int nbox = staged ? B0 * B1 * B2 : 0;
            // box element i -> volume index, with the edge clamp.  i / B is (i + 0.5) * (1/B)
            // rounded down in FP32: the quotient is below 2^13 and the nearest integer is at
            // least 0.5/B away, far more than the rounding error of the product.
            const float rB0 = 1.0f / (float)B0, rB1 = 1.0f / (float)B1;
            auto box_voxel = [&](int i) {
                const int r = (int)fmaf((float)i, rB0, 0.5f * rB0);
                const int x = min(max(lo[0] + (i - __mul24(r, B0)), 0), b.W - 1);
                const int z = (int)fmaf((float)r, rB1, 0.5f * rB1);
                const int y = min(max(lo[1] + (r - __mul24(z, B1)), 0), b.H - 1);
                const int zz = min(max(lo[2] + z, 0), b.D - 1);
                // zz*H + y < D*H < 2^31 (host check, D and H below 2^24); the whole offset is
                // 64-bit when the volume has 2^30 voxels or more (BIG)
                return (VoxelIndex)(__umul24((unsigned)zz, (unsigned)b.H) + (unsigned)y) * b.W + x;
            };
            // A box inside the volume needs no clamp: its element i = (z B1 + y) B0 + x, with
            // r = z B1 + y its row, is at offset z plane + y W + x = i + r (W - B0) + z W (H - B1)
            // from the box's first voxel -- two 24-bit multiply-adds with non-negative
            // coefficients (exact while 4 H W < 2^24, the box's cell coordinates are below 2^13
            // and the byte offsets below 2^32, i.e. without BIG).  10 full-rate VALU instructions
            // per element instead of 23 with the clamps (one of them a quarter-rate 64-bit
            // multiply-add).  At 256^3 x 720: forward 64.1 -> 61.9 ms with x and y formed
            // explicitly (15 instructions), 60.2 ms in this form, bit-identical; volume-gradient
            // adjoint 157.4 -> 155.3 ms against an incremental walk of each lane's elements i,
            // i + 256, ... (about 24 instructions per element, its uniform steps selected per
            // element), which it replaces.
            const bool inner = !BIG && staged && lo[0] >= 0 && lo[1] >= 0 && lo[2] >= 0 &&
                               hi[0] < b.W && hi[1] < b.H && hi[2] < b.D && plane < (1 << 22);
            // (unsigned arithmetic: the value is used only when inner, where it is below 2^30)
            const unsigned first = ((unsigned)lo[2] * (unsigned)b.H + (unsigned)lo[1]) * (unsigned)b.W + (unsigned)lo[0];
            const unsigned row_skip = (unsigned)(b.W - B0), plane_skip = (unsigned)b.W * (unsigned)(b.H - B1);
            auto inner_voxel = [&](int i) {      // element offset (the adjoint's staging and flush)
                const int r = (int)fmaf((float)i, rB0, 0.5f * rB0);
                const int z = (int)fmaf((float)r, rB1, 0.5f * rB1);
                return __umul24((unsigned)z, plane_skip) + __umul24((unsigned)r, row_skip) + (unsigned)i + first;
            };
            if (staged) {
                if (!BCAST) __syncthreads();
                // Box element i goes to vbox[i]; wave w stages elements c + 64 w + lane of each
                // chunk c = 0, 256, ...  A whole 64-element run is loaded straight into LDS
                // (global_load_lds_dword: lane l's dword lands at the wave-uniform LDS address +
                // 4 l), so every load of the box is in flight at once and holds no register; a
                // partial last run goes through registers.  The gradient box is zeroed alongside.
                const int wbase = __builtin_amdgcn_readfirstlane((int)(threadIdx.x & ~63u));
                if constexpr (MODE == 2) {
                    for (int i0 = wbase; i0 < nbox; i0 += SL_THREADS) {
                        const int i = i0 + lane;
                        if (TAPS) {
                            const VoxelIndex vi = inner ? (VoxelIndex)inner_voxel(i) : box_voxel(i);
                            // (without BIG the volume is below 4 GiB, so the byte offset is a 32-bit
                            // lane offset from the kernel argument's address)
                            const void* src = BIG ? (const void*)(P.vol + vi)
                                                  : (const void*)(reinterpret_cast<const char*>(P.vol) + 4u * (unsigned)vi);
                            if (i0 + 64 <= nbox)
                                __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)(vbox + i0), 4, 0, 0);
                            else if (i < nbox)
                                vbox[i] = P.vol[vi];
                        }
                        if (i < nbox) gboxu[i] = 0u;
                    }
                } else {
                    // The gathering kernels' staging, with the element's byte offset from the
                    // volume's address given by off_of (32-bit without BIG, so the loads take the
                    // kernel argument's address as their scalar base).  (The same loop for the
                    // volume-gradient adjoint measured 0.7 ms slower than the one above; this form
                    // gives the forward 0.9 ms over that one.)
                    auto stage = [&](auto off_of) {
                        for (int i0 = wbase; i0 < nbox; i0 += SL_THREADS) {
                            const int i = i0 + lane;
                            auto off = off_of(i);
                            if (i0 + 64 <= nbox) {
                                __builtin_amdgcn_global_load_lds(reinterpret_cast<const char*>(P.vol) + off,
                                                                 (__attribute__((address_space(3))) void*)(vbox + i0), 4, 0, 0);
                            } else if (i < nbox) {
                                // (an opaque copy: one 64-bit address shared with this load would
                                // cost the loads above their scalar base)
                                asm volatile("" : "+v"(off));
                                vbox[i] = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(P.vol) + off);
                            }
                        }
                    };
                    if (inner) {
                        // (the byte offset formed directly: one instruction fewer than 4 x the
                        // element offset)
                        const unsigned row4 = 4u * row_skip, plane4 = 4u * plane_skip, first4 = 4u * first;
                        stage([&](int i) {
                            const int r = (int)fmaf((float)i, rB0, 0.5f * rB0);
                            const int z = (int)fmaf((float)r, rB1, 0.5f * rB1);
                            return __umul24((unsigned)z, plane4) + __umul24((unsigned)r, row4) + 4u * (unsigned)i + first4;
                        });
                    } else {
                        stage([&](int i) {
                            const VoxelIndex vi = box_voxel(i);
                            using ByteOffset = typename std::conditional<BIG, int64_t, unsigned>::type;
                            return BIG ? (ByteOffset)(4 * (int64_t)vi) : (ByteOffset)(4u * (unsigned)vi);
                        });
                    }
                }
                // s_waitcnt vmcnt(0): the direct-to-LDS loads have landed before the barrier
                if (TAPS) __builtin_amdgcn_s_waitcnt(0x0F70);
                __syncthreads();
            }
            // advance the lane's ray through this slab (principal axis and direction as constants)
            switch (p * 2 + fwd) {
                case 0: march(std::integral_constant<int, 0>(), std::false_type(), s_lo, s_hi, lo, B0, B1, B2, staged); break;
                case 1: march(std::integral_constant<int, 0>(), std::true_type(), s_lo, s_hi, lo, B0, B1, B2, staged); break;
                case 2: march(std::integral_constant<int, 1>(), std::false_type(), s_lo, s_hi, lo, B0, B1, B2, staged); break;
                case 3: march(std::integral_constant<int, 1>(), std::true_type(), s_lo, s_hi, lo, B0, B1, B2, staged); break;
                case 4: march(std::integral_constant<int, 2>(), std::false_type(), s_lo, s_hi, lo, B0, B1, B2, staged); break;
                default: march(std::integral_constant<int, 2>(), std::true_type(), s_lo, s_hi, lo, B0, B1, B2, staged); break;
            }
            if (MODE == 2 && staged) {
                __syncthreads();
                auto flush = [&](int ival, VoxelIndex vi) {
                    if (P.gacc) {
                        const long long q = (long long)ival;
                        const int sh = E - e_tile;
                        const long long iv = sh >= 0 ? (q << min(sh, 62)) : (q >> min(-sh, 62));
                        if (iv != 0) atomicAdd(P.gacc + vi, (unsigned long long)iv);
                    } else {
                        // (a 32-bit byte offset from the argument's address, as in the staging)
                        float* dst = BIG ? P.gvol + vi
                                         : reinterpret_cast<float*>(reinterpret_cast<char*>(P.gvol) + 4u * (unsigned)vi);
                        atomicAdd(dst, (float)ival * inv_scale);
                    }
                };
                if (inner) {
                    for (int i = threadIdx.x; i < nbox; i += SL_THREADS) {
                        const int ival = (int)gboxu[i];
                        if (ival != 0) flush(ival, (VoxelIndex)inner_voxel(i));
                    }
                } else {
                    for (int i = threadIdx.x; i < nbox; i += SL_THREADS) {
                        const int ival = (int)gboxu[i];
                        if (ival != 0) flush(ival, box_voxel(i));
                    }
                }
            }
        }
    }
    // anything left (samples the slab order did not reach)
    while (alive) {
        global_step();
        t += P.step;
        alive = t <= tmax;
    }

    if (MODE == 0) {
        if (inr) P.proj[((int64_t)a * P.U + u) * P.V + v] = hit ? integral * P.step : 0.f;
        return;
    }
    if (!GEO) return;
    {
        // the constant factors of the geometry sums: upstream weight * step (1 in the calibration
        // modes, whose weight comes below) and 1/voxel size; sum over samples of grad . d
        const float sc = (MODE >= 3 ? 1.0f : gs) * P.inv_vs;
        gtmin = sc * fmaf(ray.dz, goz, fmaf(ray.dy, goy, ray.dx * gox));
        gox *= sc; goy *= sc; goz *= sc; gdx *= sc; gdy *= sc; gdz *= sc;
    }
    float loss = 0.f, resid = 0.f;
    if (MODE == 4 && inr) {
        // unit-weight sums times step: the derivatives of this ray's projection value
        resid = (hit ? integral * P.step : 0.f) - go;
        loss = 0.5f * P.resid_scale * resid * resid;
        gox *= P.step; goy *= P.step; goz *= P.step; gdx *= P.step; gdy *= P.step; gdz *= P.step;
        gtmin *= P.step;
    }
    if (MODE == 3 && inr) {
        // residual of this ray; scale the unit-weight geometry sums by d(loss)/d(proj) * step
        const float r = (hit ? integral * P.step : 0.f) - go;
        loss = 0.5f * P.resid_scale * r * r;
        const float g = P.resid_scale * r * P.step;
        gox *= g; goy *= g; goz *= g; gdx *= g; gdy *= g; gdz *= g; gtmin *= g;
    }
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (valid) {
        if (hit) {
            if (entry == 0) { gox -= gtmin / ray.dx; gdx -= gtmin * tmin / ray.dx; }
            else if (entry == 1) { goy -= gtmin / ray.dy; gdy -= gtmin * tmin / ray.dy; }
            else if (entry == 2) { goz -= gtmin / ray.dz; gdz -= gtmin * tmin / ray.dz; }
        }
        const float dd = fmaf(ray.dz, gdz, fmaf(ray.dy, gdy, ray.dx * gdx));
        const float il = 1.0f / ray.len;
        const float grx = fmaf(-ray.dx, dd, gdx) * il;
        const float gry = fmaf(-ray.dy, dd, gdy) * il;
        const float grz = fmaf(-ray.dz, dd, gdz) * il;
        const float Ox = gox - grx, Oy = goy - gry, Oz = goz - grz;
        const float dOxb = -f.sod * f.c, dOyb = -f.sod * f.s;
        const float dTxb = fmaf(-ray.up, f.s, -f.nc);
        const float dTyb = fmaf(ray.up, f.c, -f.ns);
        acc[0] = fmaf(gry, dTyb, fmaf(grx, dTxb, fmaf(Oy, dOyb, Ox * dOxb)));
        acc[1] = fmaf(grx, f.s, -(gry * f.c));
        acc[2] = fmaf(Oy + gry, f.c, (Ox + grx) * -f.s);
        acc[3] = -P.pix * fmaf(gry, f.s, grx * f.c);
        acc[4] = -P.pix * grz;
        acc[5] = Ox; acc[6] = Oy; acc[7] = Oz;
    }
    if (MODE == 4) {
        // normal equations of the tile: upper triangle of sum j j^T, sum j r, loss
        constexpr int NE = 45;
        __shared__ float red4[SL_THREADS / 64][NE];
        int idx = 0;
#pragma unroll
        for (int p = 0; p < 8; ++p)
#pragma unroll
            for (int q = p; q < 8; ++q) {
                const float s = wave_sum(acc[p] * acc[q]);
                if (lane == 0) red4[wave][idx] = s;
                ++idx;
            }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float s = wave_sum(acc[k] * resid);
            if (lane == 0) red4[wave][36 + k] = s;
        }
        {
            const float s = wave_sum(loss);
            if (lane == 0) red4[wave][44] = s;
        }
        __syncthreads();
        if (threadIdx.x < NE) {
            const int k = threadIdx.x;
            const float s = (red4[0][k] + red4[1][k]) + (red4[2][k] + red4[3][k]);
            P.partials[((int64_t)a * P.tiles_u * P.tiles_v + tu * P.tiles_v + tv) * NE + k] = s;
        }
        return;
    }
    constexpr int NP = MODE == 3 ? 9 : 8;
    __shared__ float red[SL_THREADS / 64][9];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float s = wave_sum(acc[k]);
        if (lane == 0) red[wave][k] = s;
    }
    if (MODE == 3) {
        const float s = wave_sum(loss);
        if (lane == 0) red[wave][8] = s;
    }
    __syncthreads();
    if (threadIdx.x < NP) {
        const int k = threadIdx.x;
        const float s = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
        P.partials[((int64_t)a * P.tiles_u * P.tiles_v + tu * P.tiles_v + tv) * NP + k] = s;
    }
}

template <int MODE, bool DIV, bool ZOFF, int CAPC = 0, bool GEO = true, bool BIG = false>
__global__ __launch_bounds__(SL_THREADS) __attribute__((amdgpu_waves_per_eu(MODE == 2 ? 5 : 1)))
void fp_slab_kernel(SlabArgs P) {
    // Workgroups are dispatched round-robin over the 8 XCDs; permute so that consecutive tiles of a
    // view share an XCD (and its L2).
    const int xcd = blockIdx.x & 7;
    const int tile = ((P.nblocks & 7) == 0) ? xcd * (P.nblocks >> 3) + (blockIdx.x >> 3) : (int)blockIdx.x;
    slab_tile<MODE, DIV, ZOFF, CAPC, GEO, BIG>(P, tile);
}

// Persistent calibration kernel: workgroups pull tiles of all problems from one counter.
// MODE 3 (Adam: loss gradient) or 4 (Levenberg-Marquardt: normal equations); one launch serves
// one solver, so each mode is its own kernel with its own register budget.
template <int MODE, bool BIG>
__global__ __launch_bounds__(SL_THREADS) __attribute__((amdgpu_waves_per_eu(7))) void fp_slab_batched_kernel(const CalibDev* __restrict__ probs,
                                                                     int nprob, int total_tiles,
                                                                     unsigned* __restrict__ counter) {
    __shared__ int sh_tile;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) sh_tile = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int tile = sh_tile;
        if (tile >= total_tiles) break;
        int lo = 0, hi = nprob - 1;            // last problem whose tile_offset <= tile
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (probs[mid].tile_offset <= tile) lo = mid; else hi = mid - 1;
        }
        const SlabArgs P = probs[lo].s;
        const int t = tile - probs[lo].tile_offset;
        if (probs[lo].div) slab_tile<MODE, true, false, 0, true, BIG>(P, t);
        else slab_tile<MODE, false, false, 0, true, BIG>(P, t);
    }
}

int launch_fp_slab_batched(const CalibDev* probs, int nprob, int total_tiles, int max_cap, int solver,
                           unsigned* counter, int grid, hipStream_t stream, bool big) {
    const size_t shmem = (size_t)max_cap * sizeof(float);
    if (solver == 1) {
        if (big) fp_slab_batched_kernel<4, true><<<grid, SL_THREADS, shmem, stream>>>(probs, nprob, total_tiles, counter);
        else fp_slab_batched_kernel<4, false><<<grid, SL_THREADS, shmem, stream>>>(probs, nprob, total_tiles, counter);
    } else {
        if (big) fp_slab_batched_kernel<3, true><<<grid, SL_THREADS, shmem, stream>>>(probs, nprob, total_tiles, counter);
        else fp_slab_batched_kernel<3, false><<<grid, SL_THREADS, shmem, stream>>>(probs, nprob, total_tiles, counter);
    }
    CBCT_LAUNCHED("batched calibration kernel launch");
    return CBCT_OK;
}

int launch_fp_slab(int mode, bool div, const SlabArgs& args, hipStream_t stream, bool geometry) {
    const size_t shmem = (size_t)args.cap * sizeof(float) * (mode == 2 && geometry ? 2 : 1);
    const dim3 grid((unsigned)args.nblocks), block(SL_THREADS);
    const bool zoff = args.b.offx == 0.f && args.b.offy == 0.f && args.b.offz == 0.f;
    const bool big = slab_wide_offsets(args.b.D, args.b.H, args.b.W);
    if (big) {
        // 64-bit voxel offsets: one generic variant per mode
        if (mode == 0 && div) fp_slab_kernel<0, true, false, 0, true, true><<<grid, block, shmem, stream>>>(args);
        else if (mode == 0) fp_slab_kernel<0, false, false, 0, true, true><<<grid, block, shmem, stream>>>(args);
        else if (mode == 1) fp_slab_kernel<1, false, false, 0, true, true><<<grid, block, shmem, stream>>>(args);
        else if (geometry) fp_slab_kernel<2, false, false, 0, true, true><<<grid, block, shmem, stream>>>(args);
        else fp_slab_kernel<2, false, false, 0, false, true><<<grid, block, shmem, stream>>>(args);
    } else if (mode == 2 && !geometry) {
        // volume gradient only: no taps, so neither the LDS capacity nor the offsets need to be
        // compile-time constants
        if (zoff) fp_slab_kernel<2, false, true, 0, false><<<grid, block, shmem, stream>>>(args);
        else fp_slab_kernel<2, false, false, 0, false><<<grid, block, shmem, stream>>>(args);
    } else if (mode == 0) {
        if (div) {
            if (zoff) fp_slab_kernel<0, true, true><<<grid, block, shmem, stream>>>(args);
            else fp_slab_kernel<0, true, false><<<grid, block, shmem, stream>>>(args);
        } else {
            if (zoff) fp_slab_kernel<0, false, true><<<grid, block, shmem, stream>>>(args);
            else fp_slab_kernel<0, false, false><<<grid, block, shmem, stream>>>(args);
        }
    } else if (mode == 1) {
        if (zoff) fp_slab_kernel<1, false, true><<<grid, block, shmem, stream>>>(args);
        else fp_slab_kernel<1, false, false><<<grid, block, shmem, stream>>>(args);
    } else if (zoff && args.cap == 3840) {
        fp_slab_kernel<2, false, true, 3840><<<grid, block, shmem, stream>>>(args);
    } else if (zoff && args.cap == 4096) {
        fp_slab_kernel<2, false, true, 4096><<<grid, block, shmem, stream>>>(args);
    } else if (zoff && args.cap == 2048) {
        fp_slab_kernel<2, false, true, 2048><<<grid, block, shmem, stream>>>(args);
    } else if (zoff && args.cap == 8192) {
        fp_slab_kernel<2, false, true, 8192><<<grid, block, shmem, stream>>>(args);
    } else {
        if (zoff) fp_slab_kernel<2, false, true><<<grid, block, shmem, stream>>>(args);
        else fp_slab_kernel<2, false, false><<<grid, block, shmem, stream>>>(args);
    }
    CBCT_LAUNCHED("slab projector kernel launch");
    return CBCT_OK;
}

}  // namespace cbct

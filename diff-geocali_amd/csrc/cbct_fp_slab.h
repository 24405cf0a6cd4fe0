// Launch interface of the slab-staged projector kernels (cbct_fp_slab.hip).
#pragma once

#include "cbct_common.h"

namespace cbct {

constexpr int SL_TILE = 16;
constexpr int SL_THREADS = SL_TILE * SL_TILE;

// The slab kernels' narrow variants take voxel offsets in 32 bits and volume byte offsets in 32
// unsigned bits, so they serve volumes below 2^30 voxels (4 GiB); larger ones take the variant
// with 64-bit offsets.
inline bool slab_wide_offsets(int64_t D, int64_t H, int64_t W) { return D * H * W >= (int64_t(1) << 30); }

struct SlabArgs {
    const float* vol;
    const float* grad;
    const ViewFrame* frames;
    float* proj;
    float* gvol;
    float* partials;
    Box b;
    int U, V, tiles_u, tiles_v, nblocks;
    float pix, step, vs, inv_vs;
    int T;        // slab thickness in voxel planes
    int cap;      // LDS floats per box
    int layout;   // 0: each wave owns an 8x8 pixel block; 1: waves interleave with stride 2
    float resid_scale;   // mode 3: dL/dproj = resid_scale * (proj - measured)
    // Deterministic volume gradient (mode 2): when gacc is set, contributions are added as
    // integers in units of 2^-(*gexp) to gacc (int64 two's complement), so the sum does not
    // depend on the order of the atomics; a later kernel converts gacc to gvol.
    unsigned long long* gacc;
    const int* gexp;
};

// One calibration problem of a batched launch (device resident).
struct CalibDev {
    SlabArgs s;          // mode-3 projector problem; s.grad = measured projections
    float* geom;         // [8][A] geometry, updated in place
    float* m;            // [8][A] Adam first moments
    float* v;            // [8][A] Adam second moments
    float* ggeo;         // [8][A] gradient of the loss
    float* vloss;        // [A] loss per view
    float* loss_hist;    // [nsteps] loss before each update
    int A, view_offset, tile_offset, param_offset;
    int div;             // voxel size not a power of two: the forward's division map
    int solver;          // 0: Adam on the gradient; 1: Levenberg-Marquardt on the normal equations
    float* ne;           // [A][45] normal equations of the current geometry (solver 1)
    float* lm;           // [A][LM_STATE] accepted geometry, its normal equations, damping, loss
    FrameGuard guard;    // which views can be marched (guard_frame)
};

constexpr int LM_NE = 45;               // 36 (upper triangle of J^T J) + 8 (J^T r) + loss
constexpr int LM_STATE = 8 + LM_NE + 2; // geometry, normal equations, lambda, loss

// mode 0: forward; 1: adjoint (geometry only); 2: adjoint with the volume gradient, and with the
// eight geometry gradients unless geometry is false.
int launch_fp_slab(int mode, bool div, const SlabArgs& args, hipStream_t stream, bool geometry = true);

// Mode 3 (fused forward + squared residual + geometry adjoint; solver 0) or mode 4 (normal
// equations; solver 1) over every tile of every problem, persistent workgroups fed by the work
// counter (which must be zero at launch).  Every problem of a launch uses the same solver.
// big: some problem's volume has more than 2^31 - 1 voxels (64-bit voxel offsets).
int launch_fp_slab_batched(const CalibDev* probs, int nprob, int total_tiles, int max_cap, int solver,
                           unsigned* counter, int grid, hipStream_t stream, bool big);

}  // namespace cbct

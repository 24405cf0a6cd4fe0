// Launch interface of the slab-staged projector kernels (cbct_fp_slab.hip).
#pragma once

#include "cbct_common.h"

namespace cbct {

constexpr int SL_TILE = 16;
constexpr int SL_THREADS = SL_TILE * SL_TILE;

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
};

// mode 0: forward; 1: adjoint (geometry only); 2: adjoint with the volume gradient.
int launch_fp_slab(int mode, bool div, const SlabArgs& args, hipStream_t stream);

}  // namespace cbct

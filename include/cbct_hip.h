/*
 * cbct_hip.h -- C ABI of the MI355X (gfx950) cone-beam CT operator library
 * libcbct_hip.so.
 *
 * Every entry point takes raw device pointers, int64 sizes, float scalars, a
 * caller-owned device workspace and a hipStream_t passed as void*.  Calls are
 * asynchronous on that stream and never synchronise; the caller makes the
 * stream's device current.  Outputs are allocated by the caller and fully
 * overwritten.  Return value: 0 on success, otherwise a CBCT_E* code with a
 * message available from cbct_last_error() (thread-local).
 *
 * Geometry is always passed as eight device arrays of A floats, in the order
 * angles, sid, sod, det_u0, det_v0, src_x, src_y, src_z (radians, total SDD,
 * SOD, detector-centre pixel indices, world-coordinate source offsets).
 * Geometry gradients are written to one [8][A] float array in that order.
 *
 * Citations name the reference interface each entry point replaces (paths
 * relative to the reference repository).
 */
#ifndef CBCT_HIP_H
#define CBCT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CBCT_OK = 0,
    CBCT_EINVAL = 1,   /* invalid argument (shape, scalar, null pointer, limit) */
    CBCT_EWORKSPACE = 2, /* workspace too small */
    CBCT_EHIP = 3      /* HIP runtime error (launch or device) */
};

/* Library version string, e.g. "cbct_hip 0.3.0 gfx950". */
const char* cbct_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* cbct_last_error(void);

/* ---- per-call options ----------------------------------------------------
 * No reference counterpart (the reference has one kernel per operator and no modes).
 * Kernel selection and the deterministic mode can be given per call through the *_ex
 * entry points, which then read no process-wide state: calls on different host threads or
 * streams with different options do not affect each other.  options == NULL means the
 * process-wide defaults of cbct_fp_set_tuning_op / cbct_fp_set_deterministic (which the
 * entry points without _ex always use); cbct_fp_default_options reads them.  Every setting
 * gives the same projections bit for bit; see cbct_fp_set_tuning_op for the fields. */
typedef struct {
    int use_slab;        /* 1: slab-staged kernels; 0: direct kernels (one ray per lane, global taps) */
    int slab_thickness;  /* voxel planes per staged slab, 1..64 */
    int lds_floats;      /* LDS floats of one staged box, 256..8192 */
    int layout;          /* lanes of a 16x16 ray tile: 0 = 8x8 blocks per wave, 1 = interleaved, stride 2 */
    int deterministic;   /* 1: int64 fixed-point volume gradient (projector adjoint) and projection
                            gradient (backprojection backward): bitwise reproducible */
} cbct_fp_options;

/* The process-wide defaults of operation class op (0: kernels that only gather -- forward,
 * geometry-only adjoint, calibration march; 1: adjoint with the volume gradient). */
int cbct_fp_default_options(int op, cbct_fp_options* out);

/* ---- forward projector: src_cbct_forwardprojector_cuda ----------------- */

/* Workspace bytes needed by cbct_fp_forward / cbct_fp_backward. */
size_t cbct_fp_workspace_bytes(int64_t A, int64_t U, int64_t V);

/* Fixed-step trilinear ray integrals proj[A,U,V] of volume[D,H,W].
 * Replaces forward_cuda, csrc/src_cbct_forwardprojector_cuda.cu:566-631
 * (pybind "forward", :733-744). */
int cbct_fp_forward(const float* volume, int64_t D, int64_t H, int64_t W,
                    const float* angles, const float* sid, const float* sod,
                    const float* det_u0, const float* det_v0,
                    const float* src_x, const float* src_y, const float* src_z, int64_t A,
                    float pixel_size, int64_t U, int64_t V, float step_size, float voxel_size,
                    float off_x, float off_y, float off_z,
                    float* projections, void* workspace, size_t workspace_bytes, void* stream);

/* cbct_fp_forward with per-call options (NULL: the process-wide defaults).  Replaces the same
 * forward_cuda, csrc/src_cbct_forwardprojector_cuda.cu:566-631. */
int cbct_fp_forward_ex(const float* volume, int64_t D, int64_t H, int64_t W,
                       const float* angles, const float* sid, const float* sod,
                       const float* det_u0, const float* det_v0,
                       const float* src_x, const float* src_y, const float* src_z, int64_t A,
                       float pixel_size, int64_t U, int64_t V, float step_size, float voxel_size,
                       float off_x, float off_y, float off_z,
                       float* projections, const cbct_fp_options* options,
                       void* workspace, size_t workspace_bytes, void* stream);

/* Workspace of cbct_fp_backward: cbct_fp_workspace_bytes plus, in deterministic
 * mode, 8 bytes per voxel (+256) for the integer accumulator.  Query it with the
 * deterministic setting that the call will use. */
size_t cbct_fp_backward_workspace_bytes(int64_t D, int64_t H, int64_t W, int64_t A, int64_t U, int64_t V);

/* Deterministic mode (process-wide default, off; per call: cbct_fp_options): the projector adjoint's volume
 * gradient is accumulated as int64 fixed-point values (scale 2^-E with E chosen
 * from max|grad * step|, so each sample keeps >= 40 bits), which makes the result
 * bitwise reproducible run to run; the reference's float atomics are not
 * (docs/USAGE.md:306).  Needs the slab kernels (use_slab = 1).  The same setting
 * makes cbct_bp_backward's projection gradient reproducible. */
int cbct_fp_set_deterministic(int on);
int cbct_fp_get_deterministic(void);

/* Adjoint of cbct_fp_forward for the upstream gradient grad_projection[A,U,V]:
 * grad_volume[D,H,W] (skipped and may be NULL when compute_grad_volume == 0)
 * and grad_geometry[8][A].
 * Replaces backward_cuda, csrc/src_cbct_forwardprojector_cuda.cu:634-729
 * (pybind "backward", :746-757). */
int cbct_fp_backward(const float* grad_projection, const float* volume,
                     int64_t D, int64_t H, int64_t W,
                     const float* angles, const float* sid, const float* sod,
                     const float* det_u0, const float* det_v0,
                     const float* src_x, const float* src_y, const float* src_z,
                     int64_t A, int64_t U, int64_t V,
                     float pixel_size, float step_size, float voxel_size,
                     int compute_grad_volume, float off_x, float off_y, float off_z,
                     float* grad_volume, float* grad_geometry,
                     void* workspace, size_t workspace_bytes, void* stream);

/* cbct_fp_backward with per-call options (NULL: the process-wide defaults), whose
 * workspace is cbct_fp_backward_workspace_bytes_ex with the same options.  grad_geometry may
 * be NULL: the eight geometry gradients are then not computed, and the volume-gradient
 * adjoint needs no volume taps at all (autograd calls whose geometry needs no gradient).
 * Replaces backward_cuda, csrc/src_cbct_forwardprojector_cuda.cu:634-729, including its
 * compute_grad_volume switch (source_only_forward_projector.py:110); the geometry switch has
 * no counterpart there (the reference always computes the eight sums). */
size_t cbct_fp_backward_workspace_bytes_ex(int64_t D, int64_t H, int64_t W, int64_t A, int64_t U, int64_t V,
                                           const cbct_fp_options* options);
int cbct_fp_backward_ex(const float* grad_projection, const float* volume,
                        int64_t D, int64_t H, int64_t W,
                        const float* angles, const float* sid, const float* sod,
                        const float* det_u0, const float* det_v0,
                        const float* src_x, const float* src_y, const float* src_z,
                        int64_t A, int64_t U, int64_t V,
                        float pixel_size, float step_size, float voxel_size,
                        int compute_grad_volume, float off_x, float off_y, float off_z,
                        float* grad_volume, float* grad_geometry, const cbct_fp_options* options,
                        void* workspace, size_t workspace_bytes, void* stream);

/* Allocation-free cbct_fp_backward_ex.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving or calibration
 * loops, harnesses that forbid allocation).  It takes no workspace: the workspace the call needs
 * (cbct_fp_backward_workspace_bytes_ex with the same options) must fit the library's fixed
 * device scratch of CBCT_NOALLOC_SCRATCH_BYTES (64 MiB; each GPU the library runs on holds one
 * until the process ends), otherwise the call returns CBCT_EWORKSPACE without touching the
 * device.  It calls no allocating routine.  All calls on one GPU share that GPU's scratch, so
 * calls that overlap in time on one GPU (other streams, other host threads) race: order them.
 * Every byte a call reads from the scratch is written by the same call first.  The scratch used is
 * the current device's: a stream of another device is refused with CBCT_EINVAL. */
#define CBCT_NOALLOC_SCRATCH_BYTES (64ull << 20)
int cbct_fp_backward_noalloc(const float* grad_projection, const float* volume,
                             int64_t D, int64_t H, int64_t W,
                             const float* angles, const float* sid, const float* sod,
                             const float* det_u0, const float* det_v0,
                             const float* src_x, const float* src_y, const float* src_z,
                             int64_t A, int64_t U, int64_t V,
                             float pixel_size, float step_size, float voxel_size,
                             int compute_grad_volume, float off_x, float off_y, float off_z,
                             float* grad_volume, float* grad_geometry, const cbct_fp_options* options,
                             void* stream);

/* Exact number of trilinear samples of the forward schedule (the sum over
 * rays of the loop trip counts of cbct_fp_forward), accumulated into
 * count[0] (a device uint64 the caller zeroes).  No reference counterpart:
 * it is the unit of the projectors' algorithmic traffic. */
int cbct_fp_count_samples(int64_t D, int64_t H, int64_t W,
                          const float* angles, const float* sid, const float* sod,
                          const float* det_u0, const float* det_v0,
                          const float* src_x, const float* src_y, const float* src_z, int64_t A,
                          float pixel_size, int64_t U, int64_t V, float step_size, float voxel_size,
                          float off_x, float off_y, float off_z, uint64_t* count,
                          void* workspace, size_t workspace_bytes, void* stream);

/* Process-wide tuning of cbct_fp_forward / cbct_fp_backward (no reference
 * counterpart; set by the autotuner).  use_slab selects the slab-staged kernels
 * (1) or the direct kernels (0); slab_thickness is the number of voxel planes
 * per staged slab, lds_floats the LDS capacity of one staged box, layout the
 * lane-to-pixel layout of a 16x16 ray tile (0: 8x8 blocks per wave, 1: waves
 * interleaved with stride 2).  There is one setting per operation class:
 * op 0 for the kernels that only gather (cbct_fp_forward, cbct_fp_backward
 * without the volume gradient, and the default of the calibration march), op 1
 * for cbct_fp_backward with the volume gradient; cbct_fp_set_tuning sets both
 * and cbct_fp_get_tuning reads op 0.  Every setting gives the same projections
 * bit for bit; geometry gradients are bit-identical for a given layout
 * (another layout only changes their FP32 summation order).  These are
 * process-wide defaults (reads and writes are serialised); a caller that must
 * not depend on them passes cbct_fp_options to the *_ex entry points. */
int cbct_fp_set_tuning(int use_slab, int slab_thickness, int lds_floats, int layout);
int cbct_fp_get_tuning(int* out4);
int cbct_fp_set_tuning_op(int op, int use_slab, int slab_thickness, int lds_floats, int layout);
int cbct_fp_get_tuning_op(int op, int* out4);

/* ---- weighted backprojector: src_cbct_backprojector_cuda --------------- */

/* Workspace bytes needed by cbct_bp_forward. */
size_t cbct_bp_workspace_bytes(int64_t A, int64_t D, int64_t H, int64_t W);

/* Workspace bytes needed by cbct_bp_backward: the view frames, one 8-float
 * geometry partial per workgroup and view, and, in deterministic mode
 * (cbct_fp_set_deterministic), 8 bytes per detector cell for the integer
 * accumulator.  Query it with the deterministic setting that the call will use. */
size_t cbct_bp_backward_workspace_bytes(int64_t A, int64_t U, int64_t V, int64_t D, int64_t H, int64_t W);

/* volume[D,H,W] = sum over views of FDK-weighted bilinear samples of
 * projections[A,U,V].  Replaces forward_cuda,
 * csrc/src_cbct_backprojector_cuda.cu:344-420 (pybind "forward", :525-532). */
int cbct_bp_forward(const float* projections, int64_t A, int64_t U, int64_t V,
                    const float* angles, const float* sid, const float* sod,
                    const float* det_u0, const float* det_v0,
                    const float* src_x, const float* src_y, const float* src_z,
                    float pixel_size, int64_t D, int64_t H, int64_t W, float voxel_size,
                    float off_x, float off_y, float off_z,
                    float* volume, void* workspace, size_t workspace_bytes, void* stream);

/* Backward of cbct_bp_forward for grad_volume[D,H,W]: grad_projections[A,U,V]
 * and grad_geometry[8][A].  Replaces backward_cuda,
 * csrc/src_cbct_backprojector_cuda.cu:423-522 (pybind "backward", :534-542).
 * grad_geometry is summed in a fixed order (bitwise reproducible); in
 * deterministic mode grad_projections is accumulated as int64 fixed point at a
 * per-view scale 2^-E (one sample < 2^40 units) and is bitwise reproducible too;
 * otherwise its float atomics make the last bits vary run to run. */
int cbct_bp_backward(const float* grad_volume, const float* projections,
                     int64_t A, int64_t U, int64_t V,
                     const float* angles, const float* sid, const float* sod,
                     const float* det_u0, const float* det_v0,
                     const float* src_x, const float* src_y, const float* src_z,
                     float pixel_size, int64_t D, int64_t H, int64_t W, float voxel_size,
                     float off_x, float off_y, float off_z,
                     float* grad_projections, float* grad_geometry,
                     void* workspace, size_t workspace_bytes, void* stream);

/* cbct_bp_backward with per-call options (only options->deterministic is used; NULL: the
 * process-wide default), whose workspace is cbct_bp_backward_workspace_bytes_ex.  Replaces the
 * same backward_cuda, csrc/src_cbct_backprojector_cuda.cu:423-522. */
size_t cbct_bp_backward_workspace_bytes_ex(int64_t A, int64_t U, int64_t V, int64_t D, int64_t H, int64_t W,
                                           const cbct_fp_options* options);
int cbct_bp_backward_ex(const float* grad_volume, const float* projections,
                        int64_t A, int64_t U, int64_t V,
                        const float* angles, const float* sid, const float* sod,
                        const float* det_u0, const float* det_v0,
                        const float* src_x, const float* src_y, const float* src_z,
                        float pixel_size, int64_t D, int64_t H, int64_t W, float voxel_size,
                        float off_x, float off_y, float off_z,
                        float* grad_projections, float* grad_geometry, const cbct_fp_options* options,
                        void* workspace, size_t workspace_bytes, void* stream);

/* ---- FDK ---------------------------------------------------------------- */

/* Workspace bytes needed by cbct_fdk_filter (filter only) and
 * cbct_fdk_reconstruct (filter + backprojection). */
size_t cbct_fdk_workspace_bytes(int64_t A, int64_t U, int64_t V);

/* Centred cosine preweight, Hann-windowed discrete ramp filter along U and the
 * pixel_size*sid/sod scale; sod and sid are device arrays of A values.
 * Replaces fdk_filter, cbct_backprojector/filtering.py:54-104. */
int cbct_fdk_filter(const float* projections, int64_t A, int64_t U, int64_t V,
                    float pixel_size, const float* sod, const float* sid, float alpha,
                    float* filtered, void* workspace, size_t workspace_bytes, void* stream);

/* Filter, backproject in view chunks of angle_chunk_size (each chunk summed
 * from zero, then added), scale by pi/A and clamp at zero.
 * Replaces reconstruct_fdk, cbct_backprojector/filtering.py:107-198. */
int cbct_fdk_reconstruct(const float* projections, int64_t A, int64_t U, int64_t V,
                         const float* angles, const float* sid, const float* sod,
                         const float* det_u0, const float* det_v0,
                         const float* src_x, const float* src_y, const float* src_z,
                         float pixel_size, int64_t D, int64_t H, int64_t W, float voxel_size,
                         float off_x, float off_y, float off_z,
                         float alpha, int64_t angle_chunk_size,
                         float* volume, void* workspace, size_t workspace_bytes, void* stream);

/* Extensions of the FDK helpers beyond the reference (which disclaims them,
 * cbct_backprojector/filtering.py:1-8, docs/USAGE.md:332):
 * det_u0/det_v0 (both or NULL): centre the cosine preweight on each view's
 * detector centre instead of the detector's middle; angles (or NULL): Parker
 * short-scan weights (angles increasing, angles[A-1] - angles[0] > pi; delta =
 * (range - pi)/2).  cbct_fdk_reconstruct_ex: short_scan applies Parker weights
 * and weights each view by (angles[A-1] - angles[0]) / (A - 1) instead of pi/A
 * (computed on the device; the angles must increase and span more than pi, which the
 * library does not read back to check);
 * offset_preweight uses det_u0/det_v0 in the preweight. */
int cbct_fdk_filter_ex(const float* projections, int64_t A, int64_t U, int64_t V,
                       float pixel_size, const float* sod, const float* sid, float alpha,
                       const float* det_u0, const float* det_v0, const float* angles,
                       float* filtered, void* workspace, size_t workspace_bytes, void* stream);
int cbct_fdk_reconstruct_ex(const float* projections, int64_t A, int64_t U, int64_t V,
                            const float* angles, const float* sid, const float* sod,
                            const float* det_u0, const float* det_v0,
                            const float* src_x, const float* src_y, const float* src_z,
                            float pixel_size, int64_t D, int64_t H, int64_t W, float voxel_size,
                            float off_x, float off_y, float off_z,
                            float alpha, int64_t angle_chunk_size, int short_scan, int offset_preweight,
                            float* volume, void* workspace, size_t workspace_bytes, void* stream);

/* ---- geometry calibration ------------------------------------------------
 * No reference counterpart: the reference leaves the calibration loop to user
 * code (examples/autograd.py:62-63, README.md:8).  Fits the eight per-view
 * geometry arrays of each problem by Adam to the mean squared reprojection
 * residual  L = mean over the A*U*V rays of (projection - measured)^2,  where
 * the projection is cbct_fp_forward's.  One step marches every ray once (the
 * forward value, the residual and the geometry gradient in one pass). */
typedef struct {
    const float* volume;       /* [D,H,W] known volume */
    int64_t D, H, W;
    float voxel_size;
    float off_x, off_y, off_z;
    const float* measured;     /* [A,U,V] measured projections (line integrals) */
    int64_t A, U, V;
    float pixel_size, step_size;
    float* geometry;           /* [8][A] geometry, updated in place */
    float* adam_m;             /* [8][A] Adam moments; zero them before a fresh fit */
    float* adam_v;
    float* loss_history;       /* [nsteps + 1]: L before each update and after the last */
    float* grad_out;           /* optional [8][A]: dL/dgeometry at the final geometry */
    int slab_thickness;        /* kernel tuning of this problem; 0 = library tuning */
    int lds_floats;
    int layout;
} cbct_calib_problem;

size_t cbct_calib_workspace_bytes(const cbct_calib_problem* problems, int nproblems);

/* Runs nsteps Adam steps on all problems together (uneven sizes share each
 * launch; persistent workgroups balance them).  lr8: learning rate per
 * parameter group in the geometry order; step t (counted from first_step, so a
 * fit can be continued by a later call with the same Adam moments) uses
 * lr8 * lr_decay^t and the bias corrections 1 - beta^(t+1).  loss_history[i]
 * is the loss before the i-th update of this call (each a fixed-order sum of per-view losses,
 * so bitwise reproducible).  use_graph != 0 captures all steps in one hipGraph and launches
 * it; the instantiated graph is cached by the call's sizes, workspace address and optimiser
 * settings (up to 8 graphs; cbct_calib_release_graphs frees them), so repeated calls with the
 * same key launch it again without capturing.  No part of the call synchronises the stream.
 * Graphs are captured on the current device: a stream of another device is refused with
 * CBCT_EINVAL. */
int cbct_calib_run(const cbct_calib_problem* problems, int nproblems, int nsteps, int first_step,
                   const float* lr8, float lr_decay, float beta1, float beta2, float eps, int use_graph,
                   void* workspace, size_t workspace_bytes, void* stream);

/* The same batched calibration with a Levenberg-Marquardt solver per view: each
 * iteration's fused march also forms the view's Gauss-Newton normal equations
 * (J^T J of the eight geometry derivatives of every ray's projection, J^T r
 * and the loss), and an 8x8 damped solve in FP64 gives the next geometry; an
 * iteration whose loss rises is undone for that view and retried with more
 * damping.  lambda0 = initial damping (relative to diag(J^T J)).  adam_m and
 * adam_v may be NULL.  loss_history[i] = the accepted loss after i iterations;
 * the geometry ends at the accepted iterate.  No counterpart in the reference
 * (its calibration loop is left to user code). */
/* Frees the instantiated graphs cached by cbct_calib_run / cbct_calib_run_lm (the 8 most recent)
 * and those evicted from the cache whose last launch has not yet been seen to complete (the others
 * are freed by later calls without waiting).  Call it when no calibration launched by the library
 * is still running (after synchronising the streams that ran them).  No reference counterpart. */
int cbct_calib_release_graphs(void);

size_t cbct_calib_lm_workspace_bytes(const cbct_calib_problem* problems, int nproblems);
int cbct_calib_run_lm(const cbct_calib_problem* problems, int nproblems, int nsteps, float lambda0,
                      int use_graph, void* workspace, size_t workspace_bytes, void* stream);

/* Detector data -> line integrals [A][U][V] fp32 (the projector's input format).
 * raw: [A][rows][cols] of uint16 (raw_dtype 0) or float32 (1).  transpose = 1:
 * raw is [A][V][U] (rows = V, cols = U) and is transposed to [A][U][V];
 * transpose = 0: raw is already [A][U][V] (rows = U, cols = V).  With take_log,
 * t = (raw - dark) / (flat - dark) per pixel (dark, flat: [rows][cols] fp32 in the
 * raw layout; NULL = 0 and 1) and out = -log(max(t, min_transmission)); a pixel
 * with flat <= dark gives -log(min_transmission).  Without take_log, out = raw
 * (conversion and transpose only).  Replaces the preprocessing the reference
 * leaves to the caller (README.md:97,108; docs/USAGE.md:247-251). */
int cbct_preprocess(const void* raw, int raw_dtype, const float* dark, const float* flat, int64_t A,
                    int64_t rows, int64_t cols, int transpose, float min_transmission, int take_log,
                    float* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CBCT_HIP_H */

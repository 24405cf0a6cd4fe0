/* cbct_nn.h -- BF16 MFMA kernels of the learned geometry warm start (gfx950).
 *
 * No counterpart in the reference: the learned warm start is new.  A small
 * network (a per-view 2D CNN encoder, mean pooling over views and an MLP head)
 * maps a downsampled stack of (measured, nominal reprojection) pairs to eight
 * geometry corrections per view.  Training and inference are built from the
 * entry points below; the Python side (cbct_backprojector/warmstart.py) only
 * allocates buffers and orders the calls.
 *
 * Conventions: activations are channels-last (NHWC); "bf16" buffers hold
 * bfloat16 bit patterns (uint16_t); products accumulate in FP32 on MFMA; FP32
 * values are rounded to bf16 to nearest-even.  Every call is enqueued on the
 * given stream (hipStream_t as void*) and returns 0 or a CBCT_E* code with the
 * message in cbct_last_error().  No call allocates device memory.
 */
#ifndef CBCT_NN_H
#define CBCT_NN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* C = act(op(A) . op(B) + bias), op(A) M x K, op(B) K x N.
 * trans_a = 0: A is [M][lda] (A[m*lda + k]); 1: A is [K][lda] (A[k*lda + m]).
 * trans_b = 0: B is [K][ldb] (B[k*ldb + n]); 1: B is [N][ldb] (B[n*ldb + k]).
 * act: 0 identity, 1 ReLU.  bias [N] fp32 or NULL.  Outputs: C fp32 [M][ldc]
 * and/or Cb bf16 [M][ldcb] (either may be NULL).  K must be a multiple of 8
 * unless trans_a = 1 and trans_b = 0 (both operands stored K-major); the
 * contiguous extent of each K-major operand (M for trans_a = 1, N for
 * trans_b = 0) and every leading dimension must be multiples of 8.
 * splits > 1 divides K into that many ranges (for long-K weight gradients);
 * the partial products go to the workspace and are summed in a fixed order. */
size_t cbct_nn_gemm_workspace_bytes(int64_t M, int64_t N, int splits);
int cbct_nn_gemm(int trans_a, int trans_b, int64_t M, int64_t N, int64_t K,
                 const uint16_t* A, int64_t lda, const uint16_t* B, int64_t ldb,
                 const float* bias, int act, float* C, int64_t ldc, uint16_t* Cb, int64_t ldcb,
                 int splits, void* workspace, size_t workspace_bytes, void* stream);

/* 3x3 convolution, padding 1, as a GEMM: im2col writes, for x [N][H][W][C]
 * bf16, the rows col[(n*Ho + oy)*Wo + ox][kpad] with k = (ky*3 + kx)*C + c,
 * a constant 1 at k = 9C (the bias column of the weights) and 0 beyond.
 * Ho = (H - 1)/stride + 1, Wo likewise; kpad >= 9C + 1, a multiple of 8. */
int cbct_nn_im2col(const uint16_t* x, int64_t N, int64_t H, int64_t W, int64_t C, int stride,
                   int64_t kpad, uint16_t* col, void* stream);

/* Adjoint of im2col (without the bias column): dx[n][iy][ix][c] = sum of the
 * dcol entries that im2col read from x[n][iy][ix][c].  Writes dx fp32 and/or
 * dxb bf16 (either may be NULL); each element is gathered, not accumulated. */
int cbct_nn_col2im(const float* dcol, int64_t N, int64_t H, int64_t W, int64_t C, int stride,
                   int64_t kpad, float* dx, uint16_t* dxb, void* stream);

/* dx = (y > 0) ? dy : 0 for a ReLU output y; writes dx fp32 and/or dxb bf16. */
int cbct_nn_relu_backward(const float* dy, const float* y, int64_t n, float* dx, uint16_t* dxb,
                          void* stream);

/* out[r][c] = mean over s of x[r][s][c] (fp32), and its adjoint
 * dx[r][s][c] = dout[r][c] / S written as fp32 and/or bf16. */
int cbct_nn_spatial_mean(const float* x, int64_t R, int64_t S, int64_t C, float* out, void* stream);
int cbct_nn_spatial_mean_backward(const float* dout, int64_t R, int64_t S, int64_t C, float* dx,
                                  uint16_t* dxb, void* stream);

/* View pooling: feat [P][A][C] fp32 -> h [P*A][2C] bf16 with
 * h[p,a] = (feat[p,a], mean over a' of feat[p,a']); and its adjoint
 * dfeat[p,a,c] = dh[p,a,c] + (1/A) sum over a' of dh[p,a',C+c]. */
int cbct_nn_view_concat(const float* feat, int64_t P, int64_t A, int64_t C, uint16_t* h, void* stream);
int cbct_nn_view_concat_backward(const float* dh, int64_t P, int64_t A, int64_t C, float* dfeat,
                                 void* stream);

/* out[c] = sum over r of x[r][c] (fp32, fixed order; the bias gradient). */
int cbct_nn_colsum(const float* x, int64_t R, int64_t C, float* out, void* stream);

/* Mean squared error over n values: loss[0] = mean((pred - target)^2) and
 * dpred = 2 (pred - target) / n (fp32, fixed-order reduction). */
int cbct_nn_mse(const float* pred, const float* target, int64_t n, float* loss, float* dpred, void* stream);

/* Adam on FP32 master weights w (and their bf16 copy wb, may be NULL):
 * m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
 * w -= lr (m / bc1) / (sqrt(v / bc2) + eps) + lr * weight_decay * w. */
int cbct_nn_adam(float* w, uint16_t* wb, const float* g, float* m, float* v, int64_t n, float lr,
                 float beta1, float beta2, float eps, float bc1, float bc2, float weight_decay,
                 void* stream);

/* fp32 -> bf16 (round to nearest even). */
int cbct_nn_to_bf16(const float* x, int64_t n, uint16_t* y, void* stream);

/* Network input: measured and nominal [A][U][V] fp32 projections averaged over
 * factor x factor blocks -> out [A][U/factor][V/factor][2] bf16 with channel 0
 * = scale * nominal and channel 1 = scale * diff_gain * (measured - nominal).
 * U and V must be multiples of factor. */
int cbct_nn_prepare_input(const float* measured, const float* nominal, int64_t A, int64_t U, int64_t V,
                          int factor, float scale, float diff_gain, uint16_t* out, void* stream);

/* Reprojection loss of predicted geometry corrections: per row r (a view),
 * e = pred[r] - target[r] (K values), d = J[r] e with J[r] a [Q][K] matrix
 * (the detector displacement of Q/2 fixed points per unit correction);
 * loss[0] = sum of d^2 / (R*Q) and dpred[r] = 2 J[r]^T d / (R*Q).
 * J may be shared by all rows (jstride = 0) or given per row (jstride = Q*K). */
int cbct_nn_jacobian_mse(const float* pred, const float* target, const float* J, int64_t jstride,
                         int64_t R, int64_t K, int64_t Q, float* loss, float* dpred, void* stream);

/* Mean squared difference of two fp32 arrays: out[0] = mean((a - b)^2), summed in
 * a fixed order through a workspace of cbct_nn_sqdiff_workspace_bytes() bytes
 * (the reprojection residual that decides between warm and plain start). */
size_t cbct_nn_sqdiff_workspace_bytes(void);
int cbct_nn_sqdiff_mean(const float* a, const float* b, int64_t n, float* out, void* workspace,
                        size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CBCT_NN_H */

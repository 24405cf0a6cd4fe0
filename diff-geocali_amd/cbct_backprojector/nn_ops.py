"""Torch-tensor wrappers of the warm-start network kernels (include/cbct_nn.h).

Buffers are torch tensors on the GPU: ``torch.bfloat16`` for bf16 operands
(same bit layout as the library's uint16 bf16), ``torch.float32`` otherwise.
Every function enqueues on torch's current stream and raises RuntimeError with
the library's message on a non-zero status.  There is no fallback: the
operations run in libcbct_hip.so or not at all.
"""

import torch

from . import _native

BF16 = torch.bfloat16
F32 = torch.float32


def _p(t):
    return None if t is None else t.data_ptr()


def _stream(t):
    return torch.cuda.current_stream(t.device).cuda_stream


def _call(name, *args):
    _native._check(getattr(_native.lib(), name)(*args))


def _gpu(*tensors):
    for t in tensors:
        if t is not None and t.device.type != "cuda":
            raise RuntimeError("warm-start network buffers must be GPU tensors")


def gemm(A, B, M, N, K, trans_a, trans_b, lda, ldb, C=None, Cb=None, ldc=None, ldcb=None, bias=None,
         relu=False, splits=1):
    """C / Cb = act(op(A) op(B) + bias) (see cbct_nn_gemm)."""
    _gpu(A, B, C, Cb, bias)
    lib = _native.lib()
    ws = None
    nbytes = lib.cbct_nn_gemm_workspace_bytes(M, N, int(splits)) if splits > 1 else 0
    if nbytes:
        ws = torch.empty(nbytes // 4, dtype=F32, device=A.device)
    _call("cbct_nn_gemm", int(trans_a), int(trans_b), M, N, K, A.data_ptr(), lda, B.data_ptr(), ldb, _p(bias),
          1 if relu else 0, _p(C), ldc if ldc is not None else N, _p(Cb), ldcb if ldcb is not None else N,
          int(splits), _p(ws), nbytes, _stream(A))


def im2col(x, stride, kpad):
    N, H, W, C = x.shape
    Ho, Wo = (H - 1) // stride + 1, (W - 1) // stride + 1
    col = torch.empty((N * Ho * Wo, kpad), dtype=BF16, device=x.device)
    _gpu(x)
    _call("cbct_nn_im2col", x.data_ptr(), N, H, W, C, int(stride), kpad, col.data_ptr(), _stream(x))
    return col


def col2im(dcol, shape, stride, kpad, want_bf16=False):
    N, H, W, C = shape
    dx = torch.empty((N, H, W, C), dtype=F32, device=dcol.device)
    dxb = torch.empty((N, H, W, C), dtype=BF16, device=dcol.device) if want_bf16 else None
    _gpu(dcol)
    _call("cbct_nn_col2im", dcol.data_ptr(), N, H, W, C, int(stride), kpad, dx.data_ptr(), _p(dxb), _stream(dcol))
    return dx, dxb


def relu_backward(dy, y):
    dx = torch.empty_like(dy)
    dxb = torch.empty(dy.shape, dtype=BF16, device=dy.device)
    _gpu(dy, y)
    _call("cbct_nn_relu_backward", dy.data_ptr(), y.data_ptr(), dy.numel(), dx.data_ptr(), dxb.data_ptr(),
          _stream(dy))
    return dx, dxb


def spatial_mean(x, R, S, C):
    out = torch.empty((R, C), dtype=F32, device=x.device)
    _call("cbct_nn_spatial_mean", x.data_ptr(), R, S, C, out.data_ptr(), _stream(x))
    return out


def spatial_mean_backward(dout, R, S, C):
    dx = torch.empty((R, S, C), dtype=F32, device=dout.device)
    _call("cbct_nn_spatial_mean_backward", dout.data_ptr(), R, S, C, dx.data_ptr(), None, _stream(dout))
    return dx


def view_concat(feat, P, A, C):
    """[P*A, C] fp32 -> [P*A, 2C] bf16: each view's features and their mean over the views."""
    h = torch.empty((P * A, 2 * C), dtype=BF16, device=feat.device)
    _call("cbct_nn_view_concat", feat.data_ptr(), P, A, C, h.data_ptr(), _stream(feat))
    return h


def view_concat_backward(dh, P, A, C):
    dfeat = torch.empty((P * A, C), dtype=F32, device=dh.device)
    _call("cbct_nn_view_concat_backward", dh.data_ptr(), P, A, C, dfeat.data_ptr(), _stream(dh))
    return dfeat


def colsum(x, R, C, out):
    _call("cbct_nn_colsum", x.data_ptr(), R, C, out.data_ptr(), _stream(x))


def mse(pred, target):
    loss = torch.empty(1, dtype=F32, device=pred.device)
    dpred = torch.empty_like(pred)
    _call("cbct_nn_mse", pred.data_ptr(), target.data_ptr(), pred.numel(), loss.data_ptr(), dpred.data_ptr(),
          _stream(pred))
    return loss, dpred


def adam(w, wb, g, m, v, lr, beta1, beta2, eps, bc1, bc2, weight_decay):
    _call("cbct_nn_adam", w.data_ptr(), _p(wb), g.data_ptr(), m.data_ptr(), v.data_ptr(), w.numel(), float(lr),
          float(beta1), float(beta2), float(eps), float(bc1), float(bc2), float(weight_decay), _stream(w))


def to_bf16(x, out=None):
    out = torch.empty(x.shape, dtype=BF16, device=x.device) if out is None else out
    _gpu(x)
    _call("cbct_nn_to_bf16", x.data_ptr(), x.numel(), out.data_ptr(), _stream(x))
    return out


def prepare_input(measured, nominal, factor, scale, diff_gain):
    """[A,U,V] pair -> [A, U/factor, V/factor, 2] bf16: (nominal, diff_gain * (measured - nominal)),
    block means times ``scale``."""
    A, U, V = measured.shape
    out = torch.empty((A, U // factor, V // factor, 2), dtype=BF16, device=measured.device)
    _gpu(measured, nominal)
    _call("cbct_nn_prepare_input", measured.data_ptr(), nominal.data_ptr(), A, U, V, int(factor), float(scale),
          float(diff_gain), out.data_ptr(), _stream(measured))
    return out


def jacobian_mse(pred, target, J):
    """Reprojection loss: J [Q,K] shared or [R,Q,K] per row; returns (loss [1], dpred)."""
    R, K = pred.shape
    Q = J.shape[-2]
    jstride = 0 if J.dim() == 2 else Q * K
    loss = torch.empty(1, dtype=F32, device=pred.device)
    dpred = torch.empty_like(pred)
    _gpu(pred, target, J)
    _call("cbct_nn_jacobian_mse", pred.data_ptr(), target.data_ptr(), J.data_ptr(), jstride, R, K, Q,
          loss.data_ptr(), dpred.data_ptr(), _stream(pred))
    return loss, dpred


def sqdiff_mean(a, b):
    """mean((a - b)^2) of two equally shaped fp32 tensors, as a [1] tensor."""
    if a.shape != b.shape:
        raise ValueError("sqdiff_mean needs equal shapes")
    lib = _native.lib()
    nbytes = lib.cbct_nn_sqdiff_workspace_bytes()
    ws = torch.empty(nbytes // 4, dtype=F32, device=a.device)
    out = torch.empty(1, dtype=F32, device=a.device)
    _gpu(a, b)
    _call("cbct_nn_sqdiff_mean", a.data_ptr(), b.data_ptr(), a.numel(), out.data_ptr(), ws.data_ptr(), nbytes,
          _stream(a))
    return out

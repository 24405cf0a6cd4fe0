"""FP64 CPU restatement of the learned warm-start network (test infrastructure only).

There is no reference implementation of the learned part; this module states
the network's mathematics plainly in numpy so the GPU kernels can be checked
against it.  With ``bf16=True`` it rounds values to bfloat16 at exactly the
points where the GPU path stores bf16 (GEMM operands), so the only remaining
differences are FP32-vs-FP64 accumulation and rare rounding flips; with
``bf16=False`` it is the exact FP64 network whose gradients the tests check by
central finite differences.

Layout (matches the GPU): activations NHWC; a 3x3 convolution with padding 1
is ``im2col(x) @ W.T`` where W is [C_out][kpad], column (ky*3+kx)*C_in + c
holds the tap weight, column 9*C_in the bias, columns beyond it zeros; linear
layers are ``x @ W.T + b`` with W [out][in].
"""

import numpy as np


def bf16_round(x):
    """Round to the nearest bfloat16 (ties to even) through float32, as float64 values."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = np.where(finite, (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000, u & 0xFFFF0000)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def kpad_for(c_in):
    """Columns of a conv weight: 9*C_in taps + 1 bias, rounded up to a multiple of 8."""
    return -(-(9 * c_in + 1) // 8) * 8


def out_size(n, stride):
    return (n - 1) // stride + 1


def im2col(x, stride, kpad):
    """x [N,H,W,C] -> [N*Ho*Wo, kpad]: 3x3 taps with zero padding, a bias column of ones."""
    N, H, W, C = x.shape
    Ho, Wo = out_size(H, stride), out_size(W, stride)
    xp = np.zeros((N, H + 2, W + 2, C), dtype=np.float64)
    xp[:, 1:H + 1, 1:W + 1] = x
    col = np.zeros((N, Ho, Wo, kpad), dtype=np.float64)
    for ky in range(3):
        for kx in range(3):
            tap = ky * 3 + kx
            col[..., tap * C:(tap + 1) * C] = xp[:, ky:ky + stride * (Ho - 1) + 1:stride,
                                                 kx:kx + stride * (Wo - 1) + 1:stride, :]
    col[..., 9 * C] = 1.0
    return col.reshape(N * Ho * Wo, kpad)


def col2im(dcol, shape, stride, kpad):
    """Adjoint of im2col without the bias column (scatter-add over the taps)."""
    N, H, W, C = shape
    Ho, Wo = out_size(H, stride), out_size(W, stride)
    d = dcol.reshape(N, Ho, Wo, kpad)
    dxp = np.zeros((N, H + 2, W + 2, C), dtype=np.float64)
    for ky in range(3):
        for kx in range(3):
            tap = ky * 3 + kx
            dxp[:, ky:ky + stride * (Ho - 1) + 1:stride, kx:kx + stride * (Wo - 1) + 1:stride, :] += \
                d[..., tap * C:(tap + 1) * C]
    return dxp[:, 1:H + 1, 1:W + 1]


def conv_direct(x, w, stride):
    """Brute-force 3x3 convolution (padding 1) with the bias in column 9*C: an independent
    restatement used to pin im2col."""
    N, H, W, C = x.shape
    Co = w.shape[0]
    Ho, Wo = out_size(H, stride), out_size(W, stride)
    y = np.zeros((N, Ho, Wo, Co))
    for n in range(N):
        for oy in range(Ho):
            for ox in range(Wo):
                for co in range(Co):
                    s = w[co, 9 * C]
                    for ky in range(3):
                        for kx in range(3):
                            iy, ix = oy * stride + ky - 1, ox * stride + kx - 1
                            if 0 <= iy < H and 0 <= ix < W:
                                for c in range(C):
                                    s += w[co, (ky * 3 + kx) * C + c] * x[n, iy, ix, c]
                    y[n, oy, ox, co] = s
    return y


class Config:
    """Channel plan: convs ``conv[i] -> conv[i+1]`` with stride 2, then the MLP ``fc``
    (fc[0] must be 2 * conv[-1]: per-view features and their mean over views)."""

    def __init__(self, conv=(2, 32, 64, 128, 256, 256), fc=(512, 768, 768, 8), image=128):
        self.conv = tuple(conv)
        self.fc = tuple(fc)
        self.image = int(image)
        if self.fc[0] != 2 * self.conv[-1]:
            raise ValueError("fc[0] must be twice the last conv width")

    def param_shapes(self):
        shapes = []
        for i in range(len(self.conv) - 1):
            shapes.append((f"conv{i + 1}", (self.conv[i + 1], kpad_for(self.conv[i]))))
        for i in range(len(self.fc) - 1):
            shapes.append((f"fc{i + 1}_w", (self.fc[i + 1], self.fc[i])))
            shapes.append((f"fc{i + 1}_b", (self.fc[i + 1],)))
        return shapes


def forward(cfg, params, x, P, A, bf16=True):
    """x [P*A, image, image, conv[0]] -> (out [P*A, fc[-1]], cache)."""
    r = bf16_round if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    cache = {"x": x, "P": P, "A": A, "bf16": bf16}
    h = r(x)
    nconv = len(cfg.conv) - 1
    for i in range(nconv):
        w = r(params[f"conv{i + 1}"])
        col = im2col(h, 2, w.shape[1])
        N, H, W, _ = h.shape
        Ho, Wo = out_size(H, 2), out_size(W, 2)
        y = np.maximum(col @ w.T, 0.0).reshape(N, Ho, Wo, -1)
        cache[f"conv{i + 1}"] = (h.shape, col, y)
        h = r(y)
    NI, Ho, Wo, C = cache[f"conv{nconv}"][2].shape
    y_last = cache[f"conv{nconv}"][2]
    feat = y_last.reshape(NI, Ho * Wo, C).mean(axis=1)
    f3 = feat.reshape(P, A, C)
    hcat = np.concatenate([f3, np.broadcast_to(f3.mean(axis=1, keepdims=True), f3.shape)], axis=2)
    a = r(hcat.reshape(P * A, 2 * C))
    cache["pool"] = (Ho * Wo, C)
    nfc = len(cfg.fc) - 1
    for i in range(nfc):
        w = r(params[f"fc{i + 1}_w"])
        z = a @ w.T + params[f"fc{i + 1}_b"]
        last = i == nfc - 1
        y = z if last else np.maximum(z, 0.0)
        cache[f"fc{i + 1}"] = (a, y)
        a = y if last else r(y)
    return a, cache


def mse(pred, target):
    d = pred - target
    return float((d * d).mean()), 2.0 * d / d.size


def backward(cfg, params, cache, dout):
    """Gradients of sum(dout * out) with respect to every parameter."""
    r = bf16_round if cache["bf16"] else (lambda v: np.asarray(v, dtype=np.float64))
    grads = {}
    nfc = len(cfg.fc) - 1
    d = dout
    for i in reversed(range(nfc)):
        a_in, y = cache[f"fc{i + 1}"]
        last = i == nfc - 1
        dz = d if last else np.where(y > 0, d, 0.0)
        dzb = r(dz)
        grads[f"fc{i + 1}_w"] = dzb.T @ a_in
        grads[f"fc{i + 1}_b"] = dz.sum(axis=0)
        d = dzb @ r(params[f"fc{i + 1}_w"])
    P, A = cache["P"], cache["A"]
    S, C = cache["pool"]
    dh = d.reshape(P, A, 2 * C)
    dfeat = dh[..., :C] + dh[..., C:].sum(axis=1, keepdims=True) / A
    nconv = len(cfg.conv) - 1
    shape_last = cache[f"conv{nconv}"][2].shape
    dy = np.broadcast_to((dfeat.reshape(P * A, 1, C) / S), (P * A, S, C)).reshape(shape_last)
    for i in reversed(range(nconv)):
        in_shape, col, y = cache[f"conv{i + 1}"]
        dz = np.where(y > 0, dy, 0.0).reshape(-1, y.shape[-1])
        dzb = r(dz)
        grads[f"conv{i + 1}"] = dzb.T @ col
        if i > 0:
            dcol = dzb @ r(params[f"conv{i + 1}"])
            dy = col2im(dcol, in_shape, 2, col.shape[1])
    return grads


def adam(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One Adam step (bias-corrected, decoupled weight decay), FP64."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps) - lr * wd * w
    return w, m, v

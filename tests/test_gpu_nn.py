"""Warm-start network kernels (GPU) against the FP64 oracle of oracle/nn.py.

Inputs are bf16 values, so the oracle evaluates the same products exactly in
FP64; the GPU accumulates in FP32, which bounds the GEMM error near 1e-6
relative.  For the whole network the oracle rounds to bf16 where the GPU stores
bf16, so the remaining differences are FP32 accumulation and the rare
activation whose rounding flips; the tolerances below are stated per check.
"""

import numpy as np
import pytest
import torch

import cases
from oracle import nn as onn
from cbct_backprojector import nn_ops, warmstart

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _bf(x):
    return torch.tensor(onn.bf16_round(x), dtype=torch.float32, device=DEV).to(torch.bfloat16)


def _np(t):
    return t.double().cpu().numpy()


@pytest.mark.parametrize("trans_a,trans_b,M,N,K,splits", [
    (0, 1, 100, 72, 40, 1), (0, 0, 77, 48, 96, 1), (1, 0, 64, 40, 131, 1), (1, 1, 24, 56, 64, 1),
    (1, 0, 32, 24, 5000, 7), (0, 1, 300, 264, 584, 1), (0, 0, 129, 1160, 256, 1)])
def test_gemm_matches_fp64(trans_a, trans_b, M, N, K, splits):
    rng = np.random.default_rng(M * 7 + N + K)
    a = onn.bf16_round(rng.standard_normal((M, K)))
    b = onn.bf16_round(rng.standard_normal((K, N)))
    bias = rng.standard_normal(N).astype(np.float32)
    A = _bf(a.T.copy() if trans_a else a)
    B = _bf(b.T.copy() if trans_b else b)
    lda = M if trans_a else K
    ldb = K if trans_b else N
    C = torch.empty((M, N), dtype=torch.float32, device=DEV)
    Cb = torch.empty((M, N), dtype=torch.bfloat16, device=DEV)
    nn_ops.gemm(A, B, M, N, K, trans_a, trans_b, lda, ldb, C=C, Cb=Cb, bias=torch.tensor(bias, device=DEV),
                relu=True, splits=splits)
    ref = np.maximum(a @ b + bias, 0.0)
    err = np.abs(_np(C) - ref).max() / np.abs(ref).max()
    assert err <= 2e-6
    assert np.array_equal(_np(Cb), onn.bf16_round(_np(C)))


def test_gemm_rejects_misaligned_sizes():
    A = torch.zeros((16, 12), dtype=torch.bfloat16, device=DEV)
    with pytest.raises(RuntimeError, match="multiple of 8"):
        nn_ops.gemm(A, A, 16, 16, 12, 0, 1, 12, 12, C=torch.empty((16, 16), device=DEV))


@pytest.mark.parametrize("C,H,W", [(2, 17, 12), (16, 8, 8), (8, 9, 14)])
def test_convolution_as_im2col_gemm_matches_oracle(C, H, W):
    rng = np.random.default_rng(C + H)
    x = onn.bf16_round(rng.standard_normal((3, H, W, C)))
    kp = onn.kpad_for(C)
    w = onn.bf16_round(rng.standard_normal((24, kp)))
    w[:, 9 * C + 1:] = 0.0
    col = nn_ops.im2col(_bf(x), 2, kp)
    ref_col = onn.im2col(x, 2, kp)
    assert np.array_equal(_np(col), ref_col)
    M = col.shape[0]
    y = torch.empty((M, 24), device=DEV)
    nn_ops.gemm(col, _bf(w), M, 24, kp, 0, 1, kp, kp, C=y)
    ref = onn.conv_direct(x, w, 2).reshape(M, 24)
    assert np.abs(_np(y) - ref).max() <= 2e-6 * np.abs(ref).max()


def test_col2im_matches_oracle():
    rng = np.random.default_rng(8)
    shape, C = (2, 9, 10, 8), 8
    kp = onn.kpad_for(C)
    Ho, Wo = onn.out_size(9, 2), onn.out_size(10, 2)
    d = rng.standard_normal((2 * Ho * Wo, kp)).astype(np.float32)
    dx, dxb = nn_ops.col2im(torch.tensor(d, device=DEV), shape, 2, kp, want_bf16=True)
    ref = onn.col2im(d.astype(np.float64), shape, 2, kp)
    assert np.abs(_np(dx) - ref).max() <= 1e-6 * np.abs(ref).max()
    assert np.array_equal(_np(dxb), onn.bf16_round(_np(dx)))


def test_elementwise_and_pooling_kernels():
    rng = np.random.default_rng(9)
    P, A, S, C = 2, 5, 6, 16
    y = rng.standard_normal((P * A, S, C)).astype(np.float32)
    dy = rng.standard_normal((P * A, S, C)).astype(np.float32)
    ty, tdy = torch.tensor(y, device=DEV), torch.tensor(dy, device=DEV)
    dx, dxb = nn_ops.relu_backward(tdy, ty)
    assert np.array_equal(_np(dx), np.where(y > 0, dy, 0.0))
    feat = nn_ops.spatial_mean(ty, P * A, S, C)
    assert np.abs(_np(feat) - y.astype(np.float64).mean(1)).max() <= 1e-6
    h = nn_ops.view_concat(feat, P, A, C)
    f = _np(feat).reshape(P, A, C)
    ref_h = np.concatenate([f, np.broadcast_to(f.mean(1, keepdims=True), f.shape)], 2).reshape(P * A, 2 * C)
    assert np.abs(_np(h) - onn.bf16_round(ref_h)).max() <= 2 ** -7 * np.abs(ref_h).max()
    dh = rng.standard_normal((P * A, 2 * C)).astype(np.float32)
    dfeat = nn_ops.view_concat_backward(torch.tensor(dh, device=DEV), P, A, C)
    d3 = dh.astype(np.float64).reshape(P, A, 2 * C)
    ref_df = (d3[..., :C] + d3[..., C:].sum(1, keepdims=True) / A).reshape(P * A, C)
    assert np.abs(_np(dfeat) - ref_df).max() <= 1e-6
    dsm = nn_ops.spatial_mean_backward(dfeat, P * A, S, C)
    assert np.abs(_np(dsm) - np.repeat(_np(dfeat)[:, None, :] / S, S, 1)).max() <= 1e-7
    out = torch.empty(C, device=DEV)
    nn_ops.colsum(torch.tensor(dh[:, :C].copy(), device=DEV), P * A, C, out)
    assert np.abs(_np(out) - dh[:, :C].astype(np.float64).sum(0)).max() <= 1e-5
    pred = rng.standard_normal((7, 8)).astype(np.float32)
    tgt = rng.standard_normal((7, 8)).astype(np.float32)
    loss, dpred = nn_ops.mse(torch.tensor(pred, device=DEV), torch.tensor(tgt, device=DEV))
    ref_loss, ref_d = onn.mse(pred.astype(np.float64), tgt.astype(np.float64))
    assert abs(loss.item() - ref_loss) <= 1e-6 * ref_loss
    assert np.abs(_np(dpred) - ref_d).max() <= 1e-7
    x = (rng.standard_normal(1000) * 100).astype(np.float32)
    assert np.array_equal(_np(nn_ops.to_bf16(torch.tensor(x, device=DEV))), onn.bf16_round(x))


def test_adam_kernel_matches_fp64_adam():
    rng = np.random.default_rng(10)
    n = 1000
    w = rng.standard_normal(n).astype(np.float32)
    tw, tm, tv = torch.tensor(w, device=DEV), torch.zeros(n, device=DEV), torch.zeros(n, device=DEV)
    wb = torch.empty(n, dtype=torch.bfloat16, device=DEV)
    rw, rm, rv = w.astype(np.float64), np.zeros(n), np.zeros(n)
    for t in range(1, 4):
        g = rng.standard_normal(n).astype(np.float32)
        nn_ops.adam(tw, wb, torch.tensor(g, device=DEV), tm, tv, 1e-2, 0.9, 0.999, 1e-8, 1 - 0.9 ** t,
                    1 - 0.999 ** t, 1e-4)
        rw, rm, rv = onn.adam(rw, g.astype(np.float64), rm, rv, t, 1e-2, wd=1e-4)
    assert np.abs(_np(tw) - rw).max() <= 1e-6
    assert np.array_equal(_np(wb), onn.bf16_round(_np(tw)))


def test_prepare_input_block_means():
    rng = np.random.default_rng(11)
    m = rng.random((3, 16, 12)).astype(np.float32)
    nom = rng.random((3, 16, 12)).astype(np.float32)
    x = nn_ops.prepare_input(torch.tensor(m, device=DEV), torch.tensor(nom, device=DEV), 4, 0.5, 3.0)
    ref_m = m.astype(np.float64).reshape(3, 4, 4, 3, 4).mean((2, 4)) * 0.5
    ref_n = nom.astype(np.float64).reshape(3, 4, 4, 3, 4).mean((2, 4)) * 0.5
    got = _np(x)
    assert got.shape == (3, 4, 3, 2)
    assert np.abs(got[..., 0] - ref_n).max() <= 2 ** -8 * 0.5
    assert np.abs(got[..., 1] - 3.0 * (ref_m - ref_n)).max() <= 2 ** -8 * 1.5 + 1e-6


@pytest.mark.parametrize("shared", [True, False])
def test_jacobian_mse_matches_fp64(shared):
    rng = np.random.default_rng(14)
    R, K, Q = 300, 8, 18
    pred = rng.standard_normal((R, K)).astype(np.float32)
    tgt = rng.standard_normal((R, K)).astype(np.float32)
    J = rng.standard_normal((Q, K) if shared else (R, Q, K)).astype(np.float32)
    loss, dpred = nn_ops.jacobian_mse(torch.tensor(pred, device=DEV), torch.tensor(tgt, device=DEV),
                                      torch.tensor(J, device=DEV))
    e = pred.astype(np.float64) - tgt
    Jr = np.broadcast_to(J.astype(np.float64), (R, Q, K))
    d = np.einsum("rqk,rk->rq", Jr, e)
    ref_loss = (d ** 2).sum() / (R * Q)
    ref_g = 2 * np.einsum("rqk,rq->rk", Jr, d) / (R * Q)
    assert abs(loss.item() - ref_loss) <= 1e-5 * ref_loss
    assert np.abs(_np(dpred) - ref_g).max() <= 1e-5 * np.abs(ref_g).max()


def test_sqdiff_mean_matches_fp64():
    rng = np.random.default_rng(15)
    a = rng.standard_normal(3_000_001).astype(np.float32)
    b = rng.standard_normal(3_000_001).astype(np.float32)
    got = nn_ops.sqdiff_mean(torch.tensor(a, device=DEV), torch.tensor(b, device=DEV)).item()
    ref = ((a.astype(np.float64) - b) ** 2).mean()
    assert abs(got - ref) <= 1e-5 * ref


def _small_net():
    cfg = warmstart.NetConfig(conv=(2, 16, 32, 32), fc=(64, 48, 8), image=24)
    net = warmstart.WarmStartNet(cfg, seed=3, device=DEV)
    ocfg = onn.Config(conv=cfg.conv, fc=cfg.fc, image=cfg.image)
    return cfg, ocfg, net


def test_network_forward_and_gradients_match_oracle():
    cfg, ocfg, net = _small_net()
    P, A = 2, 5
    rng = np.random.default_rng(12)
    x = onn.bf16_round(np.abs(rng.standard_normal((P * A, cfg.image, cfg.image, 2))))
    target = rng.standard_normal((P * A, 8)).astype(np.float32)
    out = net.forward(_bf(x), P, A, keep=True)
    params = net.params_numpy()
    ref_out, cache = onn.forward(ocfg, params, x, P, A, bf16=True)
    e_out = cases.rel_l2(_np(out), ref_out)
    loss, dpred = nn_ops.mse(out, torch.tensor(target, device=DEV))
    net.backward(dpred)
    _, ref_d = onn.mse(_np(out), target.astype(np.float64))
    ref_g = onn.backward(ocfg, params, cache, ref_d)
    got_g = net.grads_numpy()
    errs = {k: cases.rel_l2(got_g[k], ref_g[k]) for k in ref_g}
    print(f"network forward rel L2 {e_out:.2e}; gradients", {k: f"{v:.1e}" for k, v in errs.items()})
    assert e_out <= 2e-3
    assert max(errs.values()) <= 1e-2


def test_training_steps_follow_fp64_adam_on_oracle_gradients():
    """Two Adam steps of the GPU network equal two FP64 Adam steps driven by the oracle's
    (bf16-mirrored) gradients, to the tolerance of the gradient comparison."""
    cfg, ocfg, net = _small_net()
    P, A = 1, 6
    rng = np.random.default_rng(13)
    x = onn.bf16_round(np.abs(rng.standard_normal((P * A, cfg.image, cfg.image, 2))))
    target = rng.standard_normal((P * A, 8)).astype(np.float32)
    params = net.params_numpy()
    m = {k: np.zeros_like(v) for k, v in params.items()}
    v = {k: np.zeros_like(v) for k, v in params.items()}
    lr = 1e-3
    for t in (1, 2):
        out, cache = onn.forward(ocfg, params, x, P, A, bf16=True)
        _, d = onn.mse(out, target.astype(np.float64))
        g = onn.backward(ocfg, params, cache, d)
        for k in params:
            params[k], m[k], v[k] = onn.adam(params[k], g[k], m[k], v[k], t, lr)
        net.train_step(_bf(x), torch.tensor(target, device=DEV), P, A, lr)
    got = net.params_numpy()
    moved = max(np.abs(got[k] - params[k]).max() for k in params) / lr
    print(f"max |GPU - FP64| parameter difference after 2 steps: {moved:.3f} lr")
    assert moved <= 0.5

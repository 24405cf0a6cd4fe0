"""The FP64 oracle of the warm-start network, pinned on the CPU: bf16 rounding against
torch's conversion, im2col-convolution against a brute-force convolution, col2im as the
adjoint of im2col, the backward pass against central finite differences, and the
default network's size."""

import numpy as np
import pytest
import torch

from oracle import nn as onn


def test_bf16_round_matches_torch_conversion_and_ties_to_even():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(4000) * 10.0 ** rng.integers(-30, 30, 4000),
                        [0.0, -0.0, 1.0, 3.0e38, -1e-40, 65504.0]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(onn.bf16_round(x), ref)
    # 1 + 2^-8 lies halfway between 1 and 1 + 2^-7: ties go to the even mantissa (1.0)
    assert onn.bf16_round(np.float32(1 + 2 ** -8)) == 1.0
    assert onn.bf16_round(np.float32(1 + 3 * 2 ** -8)) == 1 + 2 ** -6


@pytest.mark.parametrize("stride,C", [(1, 2), (2, 3), (2, 8)])
def test_im2col_convolution_equals_direct_convolution(stride, C):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 7, 6, C))
    kp = onn.kpad_for(C)
    w = rng.standard_normal((5, kp))
    w[:, 9 * C + 1:] = 0.0
    y = (onn.im2col(x, stride, kp) @ w.T).reshape(2, onn.out_size(7, stride), onn.out_size(6, stride), 5)
    assert np.allclose(y, onn.conv_direct(x, w, stride), rtol=1e-12, atol=1e-12)


def test_col2im_is_the_adjoint_of_im2col():
    rng = np.random.default_rng(2)
    shape, stride, C = (2, 9, 8, 4), 2, 4
    kp = onn.kpad_for(C)
    x = rng.standard_normal(shape)
    col = onn.im2col(x, stride, kp)
    d = rng.standard_normal(col.shape)
    d[:, 9 * C:] = 0.0                      # the bias column is not a function of x
    lhs = float((col * d).sum())
    rhs = float((x * onn.col2im(d, shape, stride, kp)).sum())
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def _small():
    cfg = onn.Config(conv=(2, 8, 16), fc=(32, 16, 8), image=8)
    rng = np.random.default_rng(3)
    params = {}
    for name, shape in cfg.param_shapes():
        params[name] = rng.standard_normal(shape) * 0.3
    for i, c in enumerate(cfg.conv[:-1]):
        params[f"conv{i + 1}"][:, 9 * c + 1:] = 0.0
    x = np.abs(rng.standard_normal((2 * 3, 8, 8, 2)))
    return cfg, params, x


def test_oracle_backward_matches_central_differences():
    cfg, params, x = _small()
    P, A = 2, 3
    target = np.random.default_rng(4).standard_normal((P * A, 8))
    out, cache = onn.forward(cfg, params, x, P, A, bf16=False)
    _, dout = onn.mse(out, target)
    grads = onn.backward(cfg, params, cache, dout)
    rng = np.random.default_rng(5)
    h = 1e-6
    for name in params:
        flat = params[name].reshape(-1)
        idx = [i for i in rng.permutation(flat.size) if not (name.startswith("conv") and params[name].shape[1] > 0
                                                             and (i % params[name].shape[1]) > 9 * cfg.conv[int(name[4:]) - 1])][:6]
        for i in idx:
            old = flat[i]
            flat[i] = old + h
            lp, _ = onn.mse(onn.forward(cfg, params, x, P, A, bf16=False)[0], target)
            flat[i] = old - h
            lm, _ = onn.mse(onn.forward(cfg, params, x, P, A, bf16=False)[0], target)
            flat[i] = old
            fd = (lp - lm) / (2 * h)
            g = grads[name].reshape(-1)[i]
            assert abs(fd - g) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (name, i, fd, g)


def test_adam_first_step_moves_by_lr_against_the_gradient_sign():
    w = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    w1, m, v = onn.adam(w, g, np.zeros(3), np.zeros(3), 1, lr=0.01, eps=1e-12)
    assert np.allclose(w1, w - 0.01 * np.sign(g), rtol=0, atol=1e-9)


def test_default_network_has_about_two_million_parameters():
    from cbct_backprojector.warmstart import NetConfig
    cfg = NetConfig()
    n = sum(int(np.prod(s)) for _, s in cfg.param_shapes())
    assert 1.8e6 <= n <= 2.2e6
    assert [s for _, s in cfg.param_shapes()] == [s for _, s in onn.Config().param_shapes()]
    assert cfg.feature_size() == 4

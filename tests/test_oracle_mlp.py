"""The CPU restatement of the warm-start perceptron (oracle/mlp.py), pinned by FP64
central finite differences and by the BF16 rounding rule."""

import numpy as np
import pytest

from oracle import mlp


def test_parameter_count_and_layout():
    assert mlp.NPARAMS == 16 * 128 + 128 + 128 * 128 + 128 + 128 * 3 + 3 == 19075
    p = np.arange(mlp.NPARAMS, dtype=np.float64)
    assert np.array_equal(mlp.pack(mlp.unpack(p)), p)
    P = mlp.unpack(p)
    assert P["W1"][0, 1] == 1 and P["b1"][0] == 2048 and P["W2"][0, 0] == 2048 + 128


def test_bf16_rounding_rule():
    # representable values are exact; ties go to even; halfway above rounds up
    assert mlp.bf16(np.float32(1.5))[()] == 1.5
    one = np.float32(1.0).view(np.uint32)
    tie_even = np.array([one + 0x8000], dtype=np.uint32).view(np.float32)   # 1 + 2^-8: tie, keep even 1.0
    tie_odd = np.array([one + 0x18000], dtype=np.uint32).view(np.float32)   # (1 + 2^-7) + 2^-8: tie, up
    above = np.array([one + 0x8001], dtype=np.uint32).view(np.float32)
    assert mlp.bf16(tie_even)[0] == 1.0
    assert mlp.bf16(tie_odd)[0] == 1.0 + 2.0 ** -6
    assert mlp.bf16(above)[0] == 1.0 + 2.0 ** -7
    assert np.isnan(mlp.bf16(np.float32(np.nan)))
    x = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    assert np.all(np.abs(mlp.bf16(x) - x) <= np.abs(x) * 2.0 ** -8)


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(3)
    p = mlp.init(seed=1).astype(np.float64)
    p[mlp.NPARAMS - 3:] = 0.1  # b3
    X = rng.normal(size=(24, 16))
    T = rng.normal(size=(24, 3))
    loss, g = mlp.loss_and_grad(p, X, T)
    idx = rng.choice(mlp.NPARAMS, size=60, replace=False)
    idx = np.r_[idx, 2048 + 5, 2048 + 128 + 16384 + 7, mlp.NPARAMS - 1]  # a bias of each layer
    h = 1e-6
    for i in idx:
        e = np.zeros(mlp.NPARAMS)
        e[i] = h
        fd = (mlp.loss_and_grad(p + e, X, T)[0] - mlp.loss_and_grad(p - e, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # zero error gives zero gradient
    y = mlp.predict(p, X)  # network output + the transported impulse of features 8-10
    l0, g0 = mlp.loss_and_grad(p, X, y)
    assert l0 == 0.0 and np.all(g0 == 0.0)


def test_bf16_emulation_stays_close_to_exact():
    rng = np.random.default_rng(5)
    p = mlp.init(seed=2)
    X = rng.normal(size=(64, 16))
    T = rng.normal(size=(64, 3))
    l_ex, g_ex = mlp.loss_and_grad(p, X, T)
    l_bf, g_bf = mlp.loss_and_grad(p, X, T, emulate_bf16=True)
    assert abs(l_bf - l_ex) <= 2e-2 * l_ex
    assert np.linalg.norm(g_bf - g_ex) <= 5e-2 * np.linalg.norm(g_ex)


def test_zero_output_layer_reproduces_the_transported_impulse():
    """The prediction is the network output plus features 8-10 (the transported cached
    impulse): with the output layer zero it is exactly those features."""
    rng = np.random.default_rng(9)
    p = mlp.init(seed=4)
    P = mlp.unpack(p)
    P["W3"][:] = 0.0
    P["b3"][:] = 0.0
    p = mlp.pack(P)
    X = rng.normal(size=(32, 16))
    assert np.array_equal(mlp.predict(p, X), X[:, 8:11])
    loss, g = mlp.loss_and_grad(p, X, X[:, 8:11])
    assert loss == 0.0 and np.all(g == 0.0)

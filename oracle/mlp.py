"""CPU restatement of the warm-start perceptron (16-128-128-3, ReLU, mean squared error).

TEST INFRASTRUCTURE ONLY: the checker the device's BF16 forward/backward kernels are
compared against (tests/ only). The perceptron has no counterpart in the reference; its
gradients are pinned here by central finite differences in FP64 (tests/test_oracle_mlp.py),
and `emulate_bf16=True` rounds at the points where the device kernels hold BF16 values
(inputs, weights, hidden activations, output and hidden gradients; FP32/FP64 sums).

Parameter vector layout (FP32): W1 [16][128], b1 [128], W2 [128][128], b2 [128],
W3 [128][3], b3 [3]; y = W3^T relu(W2^T relu(W1^T x + b1) + b2) + b3.
"""

import numpy as np

IN, HID, OUT = 16, 128, 3
SHAPES = [("W1", (IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, OUT)), ("b3", (OUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, k = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[k:k + n], dtype=np.float64).reshape(shape)
        k += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.atleast_1d(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = (r.astype(np.uint32)).view(np.float32).astype(np.float64)
    out[np.isnan(f)] = np.nan
    return out.reshape(np.shape(x))


def init(seed=0):
    """He-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / shape[0])
            d[name] = rng.uniform(-lim, lim, size=shape)
        else:
            d[name] = np.zeros(shape)
    return pack(d).astype(np.float32)


def forward(p, X, emulate_bf16=False):
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    P = unpack(p)
    x = r(X)
    h1 = r(np.maximum(x @ r(P["W1"]) + P["b1"], 0.0))
    h2 = r(np.maximum(h1 @ r(P["W2"]) + P["b2"], 0.0))
    y = h2 @ r(P["W3"]) + P["b3"]
    return x, h1, h2, y


BASE = 8  # the network predicts a correction to features 8-10 (the transported cached impulse)


def predict(p, X, emulate_bf16=False):
    """Predicted (scaled) impulses: network output plus features 8-10."""
    return forward(p, X, emulate_bf16)[3] + np.asarray(X, dtype=np.float64)[:, BASE:BASE + OUT]


def loss_and_grad(p, X, T, emulate_bf16=False):
    """Mean over samples and outputs of (y + x[8:11] - t)^2, and its gradient (parameter
    layout)."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    P = unpack(p)
    x, h1, h2, y = forward(p, X, emulate_bf16)
    n = X.shape[0]
    d = y + np.asarray(X, dtype=np.float64)[:, BASE:BASE + OUT] - np.asarray(T, dtype=np.float64)
    loss = float((d * d).sum() / (n * OUT))
    gy = r(2.0 * d / (n * OUT))
    g2 = r((gy @ r(P["W3"]).T) * (h2 > 0))
    g1 = r((g2 @ r(P["W2"]).T) * (h1 > 0))
    G = {"W3": h2.T @ gy, "b3": gy.sum(0), "W2": h1.T @ g2, "b2": g2.sum(0), "W1": x.T @ g1, "b1": g1.sum(0)}
    return loss, pack(G)

"""Learned warm start: the perceptron's parameter layout and initialisation.

The model (16-128-128-3, ReLU) runs and trains in libscx.so (BF16 MFMA kernels); this
module only creates an initial FP32 parameter vector in the layout of include/scx.h
(SCX_MLP_PARAMS): W1 [16][128], b1 [128], W2 [128][128], b2 [128], W3 [128][3], b3 [3].
"""

import numpy as np

SHAPES = [("W1", (16, 128)), ("b1", (128,)), ("W2", (128, 128)), ("b2", (128,)), ("W3", (128, 3)), ("b3", (3,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def init_params(seed=0):
    """He-uniform hidden weights (limit sqrt(6 / fan_in)), zero output weights and biases:
    the prediction starts as the transported cached impulse (the network adds a learned
    correction to it)."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        if name == "W3":
            parts.append(np.zeros(int(np.prod(shape))))
        elif name.startswith("W"):
            lim = np.sqrt(6.0 / shape[0])
            parts.append(rng.uniform(-lim, lim, size=shape).ravel())
        else:
            parts.append(np.zeros(int(np.prod(shape))))
    return np.concatenate(parts).astype(np.float32)

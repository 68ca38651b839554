"""The learned warm start's perceptron on the device: BF16 MFMA forward/backward
gradients against the CPU restatement (oracle/mlp.py, BF16 rounding emulated), and Adam
training on the device."""

import numpy as np
import pytest

import scisim_amd as scx
from oracle import mlp

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = scx.Context(3, 1000, max_planes=8)
    yield c
    c.close()


@pytest.mark.parametrize("n", [16, 200, 1000])
def test_device_gradients_match_oracle(ctx, n):
    rng = np.random.default_rng(n)
    p = mlp.init(seed=7)
    p[mlp.NPARAMS - 3:] = np.float32(0.05)
    ctx.set_warm_model(p)
    assert np.array_equal(ctx.warm_model(), p)
    X = rng.normal(size=(n, 16)).astype(np.float32)
    T = rng.normal(size=(n, 3)).astype(np.float32)
    g_dev, loss_dev = ctx.warm_model_gradients(X, T)
    loss_ref, g_ref = mlp.loss_and_grad(p, X, T, emulate_bf16=True)
    loss_ex, g_ex = mlp.loss_and_grad(p, X, T)
    _say("n", n, "loss dev", loss_dev, "oracle bf16", loss_ref, "exact", loss_ex)
    assert abs(loss_dev - loss_ref) <= 2e-3 * loss_ref
    k = 0
    for name, shape in mlp.SHAPES:
        m = int(np.prod(shape))
        a, b = g_dev[k:k + m].astype(np.float64), g_ref[k:k + m]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        _say(name, "relative gradient error", err)
        assert err <= 2e-2, name
        k += m
    # and the device gradient points the way the exact one does
    assert g_dev.astype(np.float64) @ g_ex >= 0.98 * np.linalg.norm(g_dev) * np.linalg.norm(g_ex)


def test_zero_network_predicts_the_transported_impulse(ctx):
    """With the output layer zero the device prediction is features 8-10 exactly: zero loss
    and zero gradient against those targets."""
    rng = np.random.default_rng(2)
    P = mlp.unpack(mlp.init(seed=6))
    P["W3"][:] = 0.0
    P["b3"][:] = 0.0
    ctx.set_warm_model(mlp.pack(P).astype(np.float32))
    X = rng.normal(size=(500, 16)).astype(np.float32)
    g, loss = ctx.warm_model_gradients(X, X[:, 8:11].copy())
    assert loss == 0.0 and np.all(g == 0.0)


def test_device_training_fits_a_smooth_map(ctx):
    """Adam on the device drives the loss of a learnable map well below its start."""
    rng = np.random.default_rng(11)
    n = 8192
    X = rng.normal(size=(n, 16)).astype(np.float32)
    A = rng.normal(size=(16, 3)) / 4.0
    T = (np.tanh(X @ A) + 0.1 * (X[:, :3] ** 2)).astype(np.float32)
    p = mlp.init(seed=3)
    ctx.set_warm_model(p)
    _, loss0 = ctx.warm_model_gradients(X[:1024], T[:1024])
    loss = ctx.train_warm_model(X, T, epochs=20, batch=512, lr=2e-3, seed=5)
    _say("loss before", loss0, "after", loss)
    assert np.isfinite(loss) and loss < 0.1 * loss0
    # the trained parameters are what the device reports, and the oracle agrees on the loss
    q = ctx.warm_model()
    l_ref, _ = mlp.loss_and_grad(q, X[:512], T[:512], emulate_bf16=True)
    assert abs(l_ref - loss) <= 5e-3 * max(l_ref, 1e-12) + 1e-6


def test_learned_warm_start_cuts_sweeps_on_a_settled_pile():
    """On a settled pile, the perceptron trained on the device on the pile's own
    (feature, converged impulse) samples warm-starts the solve: every step still reaches
    the tolerance, with far fewer sweeps than cold starts from the same states."""
    from scisim_amd import scenes, warmstart
    s = scenes.spheres_pile(n=3000, seed=4)
    ctx = scx.Context(3, 3000, max_planes=8)
    ctx.set_planes(s["px"], s["pn"])
    ctx.set_bodies(s["q"], s["v"], s["r"], s["m"], inertia=s["inertia"])
    base = dict(dt=0.005, g=[0.0, -9.81, 0.0], mu=0.4, tol=1e-6, max_iters=5000, integrator=1)

    def params(mode):
        return scx.step_params(precision=scx.FP32_POLISH, warm_start=mode, **base)

    for _ in range(300):  # settle under the transported cache
        ctx.step(params(scx.WARM_TRANSPORT))
    Xs, Ts = [], []
    for _ in range(30):
        ctx.step(params(scx.WARM_TRANSPORT))
        X, T = ctx.samples()
        Xs.append(X)
        Ts.append(T)
    X, T = np.concatenate(Xs), np.concatenate(Ts)
    ctx.set_warm_model(warmstart.init_params(seed=2))
    loss = ctx.train_warm_model(X, T, epochs=20, batch=1024, lr=1e-3, seed=3)
    q, v, w = ctx.get_state()
    sweeps = {}
    for name, mode in (("cold", scx.WARM_NONE), ("learned", scx.WARM_LEARNED)):
        ctx.set_bodies(q, v, s["r"], s["m"], omega=w, inertia=s["inertia"])
        ctx.step(params(scx.WARM_TRANSPORT))
        sw = []
        for _ in range(8):
            st = ctx.step(params(mode))
            assert st["succeeded"], (name, st)
            sw.append(st["sweeps"])
        sweeps[name] = sw
    ctx.close()
    _say("samples", X.shape[0], "loss", loss, "sweeps", sweeps)
    assert sum(sweeps["learned"]) * 2 <= sum(sweeps["cold"])


def test_learned_prediction_beats_the_transported_impulse_on_held_out_steps():
    """The learned warm start against the transported cache it starts from, on steps the
    perceptron never saw: a settled pile of 20,000 spheres yields (feature, converged
    impulse) samples of 16 transported-cache steps; the first 12 train the model on the
    device, the last 4 are held out. The device's own mean-squared error of the prediction
    (scx_warm_model_gradients) on the held-out samples must be lower with the trained model
    than with the zero network, whose prediction is the transported impulse itself: over
    all contacts, and by a wider margin on the contacts new in their step (no cached
    impulse, feature 11 = 0), where the transported guess is zero."""
    from scisim_amd import scenes, warmstart
    n = 20000
    s = scenes.spheres_pile(n=n, seed=6)
    ctx = scx.Context(3, n, max_planes=8)
    ctx.set_planes(s["px"], s["pn"])
    ctx.set_bodies(s["q"], s["v"], s["r"], s["m"], inertia=s["inertia"])
    base = dict(dt=0.005, g=[0.0, -9.81, 0.0], mu=0.4, tol=1e-6, integrator=1)
    settle = scx.step_params(precision=scx.FP32_POLISH, warm_start=scx.WARM_TRANSPORT, max_iters=1000, **base)
    for _ in range(300):
        ctx.step(settle)
    tp = scx.step_params(precision=scx.FP32_POLISH, warm_start=scx.WARM_TRANSPORT, max_iters=50000, **base)
    Xs, Ts = [], []
    for _ in range(16):
        st = ctx.step(tp)
        assert st["succeeded"], st
        X, T = ctx.samples()
        Xs.append(X)
        Ts.append(T)
    Xtr, Ttr = np.concatenate(Xs[:12]), np.concatenate(Ts[:12])
    Xte, Tte = np.concatenate(Xs[12:]), np.concatenate(Ts[12:])
    new = Xte[:, 11] == 0

    def mse(X, T):
        return ctx.warm_model_gradients(X, T)[1]

    ctx.set_warm_model(warmstart.init_params(seed=2))
    base_all, base_new = mse(Xte, Tte), mse(Xte[new], Tte[new])
    ctx.train_warm_model(Xtr, Ttr, epochs=12, batch=2048, lr=1e-3, seed=3)
    got_all, got_new = mse(Xte, Tte), mse(Xte[new], Tte[new])
    ctx.close()
    _say("train", Xtr.shape[0], "held out", Xte.shape[0], "new", int(new.sum()), "mse transported / learned: all",
         base_all, got_all, "new", base_new, got_new)
    assert new.sum() >= 100
    assert got_all < base_all
    assert got_new < 0.8 * base_new

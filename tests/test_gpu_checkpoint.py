"""Bit-identical resume, mirroring the reference's serialization test
(assets/ball2d/shell_scripts/execute_serialization_test.sh:77-132): a full run writes a
checkpoint at an intermediate step and one at the end; a fresh context resumed from the
intermediate checkpoint runs to the end, and its final checkpoint image, state and last
step's forces must equal the full run's byte for byte (the test's `diff` of the serial
files and `h5diff` of the configuration and force files)."""

import sys

import numpy as np
import pytest

import scisim_amd as scx
from scisim_amd import scenes, warmstart

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(ctx, params, k0, k1):
    st = None
    for k in range(k0, k1):
        p = params(k)
        st = ctx.step(p)
    return st


def _final(ctx):
    q, v, w = ctx.get_state()
    c = ctx.contacts()
    return q, v, w, c


def _assert_identical(a, b):
    qa, va, wa, ca = a
    qb, vb, wb, cb = b
    assert np.array_equal(qa, qb) and np.array_equal(va, vb)
    if wa is not None:
        assert np.array_equal(wa, wb)
    for k in ("a", "b", "plane", "basis", "alpha", "beta", "colour"):
        assert np.array_equal(ca[k], cb[k]), k


def _resume_case(make_ctx, params, total, resume_at, tag):
    ctx = make_ctx()
    try:
        _run(ctx, params, 0, resume_at)
        mid = ctx.save_checkpoint()
        _run(ctx, params, resume_at, total)
        full = _final(ctx)
        end_full = ctx.save_checkpoint()
    finally:
        ctx.close()
    info = scx.checkpoint_info(mid)
    _say(tag, "checkpoint at step", resume_at, len(mid), "bytes", info)
    # the resumed context starts empty and smaller than the scene (the load grows it)
    dim = info["dim"]
    ctx2 = scx.Context(dim, 16, max_planes=4)
    try:
        ctx2.load_checkpoint(mid)
        assert ctx2.save_checkpoint() == mid
        st = _run(ctx2, params, resume_at, total)
        resumed = _final(ctx2)
        end_resumed = ctx2.save_checkpoint()
    finally:
        ctx2.close()
    _say(tag, "resumed to step", total, "contacts", st["ncontacts"], "sweeps", st["sweeps"])
    assert end_resumed == end_full
    _assert_identical(full, resumed)
    return info


def test_ball2d_resume_with_the_impulse_cache_is_bit_identical():
    s = scenes.ball2d_lattice(m=48, side=7.5, r=0.05, vstd=0.1, seed=3)
    prm = dict(dt=0.01, g=[0.0, -9.81], mu=0.5, CoR=0.0, tol=1e-6, max_iters=5000, eval_every=1, integrator=0,
               precision=scx.FP32_POLISH, warm_start=scx.WARM_CACHE)

    def make():
        ctx = scx.Context(2, s["r"].size, max_planes=4)
        ctx.set_planes(s["px"], s["pn"])
        ctx.set_bodies(s["q"], s["v"], s["r"], s["m"])
        return ctx

    info = _resume_case(make, lambda k: scx.step_params(iteration=k + 1, **prm), 60, 35, "ball2d")
    assert info["cache_entries"] > 0


def test_rigidbody3d_resume_symplectic_euler_learned_is_bit_identical():
    """The reference's 3D warm-start path (symplectic Euler staggering with the
    transported cache) with the learned warm start: model and Adam moments travel too,
    and training continued after the resume matches."""
    s = scenes.spheres_pile(n=3000, rmed=0.05, sigma=0.25, seed=7)
    base = dict(dt=0.005, g=[0.0, -9.81, 0.0], mu=0.4, CoR=0.0, tol=1e-6, max_iters=3000, eval_every=1,
                integrator=1, precision=scx.FP32_POLISH, staggering=scx.STAGGER_SYMPLECTIC_EULER)

    def make():
        ctx = scx.Context(3, s["r"].size, max_planes=8)
        ctx.set_planes(s["px"], s["pn"])
        ctx.set_bodies(s["q"], s["v"], s["r"], s["m"], inertia=s["inertia"])
        tp = scx.step_params(warm_start=scx.WARM_TRANSPORT, **base)
        Xs, Ts = [], []
        for _ in range(6):
            ctx.step(tp)
            X, T = ctx.samples()
            Xs.append(X)
            Ts.append(T)
        ctx.set_warm_model(warmstart.init_params(seed=2))
        ctx.train_warm_model(np.concatenate(Xs), np.concatenate(Ts), epochs=2, batch=1024, lr=1e-3, seed=2)
        return ctx

    info = _resume_case(make, lambda k: scx.step_params(warm_start=scx.WARM_LEARNED, **base), 14, 6, "rigidbody3d")
    assert info["has_model"] and info["cache_entries"] > 0


def test_ensemble_resume_is_bit_identical():
    s, body_off, plane_off, mus, counts = scenes.ensemble_piles(nscenes=6, nmin=300, nmax=1500, seed=11)
    prm = dict(dt=0.005, g=[0.0, -9.81, 0.0], mu=0.5, CoR=0.0, tol=1e-6, max_iters=2000, eval_every=1, integrator=1,
               precision=scx.FP32_POLISH, warm_start=scx.WARM_TRANSPORT)

    def make():
        ctx = scx.Context(3, s["r"].size, max_planes=64)
        ctx.set_planes(s["px"], s["pn"])
        ctx.set_bodies(s["q"], s["v"], s["r"], s["m"], inertia=s.get("inertia"))
        ctx.set_scenes(body_off, plane_off, mus)
        return ctx

    info = _resume_case(make, lambda k: scx.step_params(**prm), 12, 5, "ensemble")
    assert info["scenes"] == 6


def test_load_refuses_a_checkpoint_of_the_other_dimension():
    s = scenes.ball2d_lattice(m=8, side=2.0, r=0.05, vstd=0.1, seed=1)
    ctx = scx.Context(2, s["r"].size, max_planes=4)
    try:
        ctx.set_planes(s["px"], s["pn"])
        ctx.set_bodies(s["q"], s["v"], s["r"], s["m"])
        img = ctx.save_checkpoint()
    finally:
        ctx.close()
    ctx3 = scx.Context(3, 16, max_planes=4)
    try:
        with pytest.raises(scx.ScxError):
            ctx3.load_checkpoint(img)
    finally:
        ctx3.close()

"""The FP32 phase's shortcut past full residual evaluations.

A full evaluation of the FP32 sweeps also lists the contacts whose residual is within a
factor 10 of its maximum (at least the stopping value). While workgroup 0 finds one of
them still at or above the stopping value after a sweep, the maximum over all contacts is
too, so that sweep's full evaluation can only say "go on" and is skipped; once none is,
the full evaluation runs and decides as before. The iterates, the sweep counts and the
residuals must therefore be those of a full evaluation after every sweep
(SCX_VIOLATOR_SKIP=0) bit for bit, and on these scenes some evaluations must actually be
skipped (SCX_VIOLATOR_SKIP=2 fails a step whose FP32 phase ran 64 sweeps or more without
skipping one). Skipping stops at sweep 768 of a launch, so that the FP32 phase's
stagnation test (from sweep 1024 on) sees the same best residuals: the cold steps, whose
FP32 phases run into it, must match as well. The evaluation counters (scx_get_sweep_counters)
must add up: one evaluation before the first sweep of each phase and one after every sweep
(eval_every = 1), less the skipped ones, which the run without skipping performs in full.
"""

import sys

import numpy as np
import pytest

import scisim_amd as scx
from scisim_amd import scenes

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def circles():
    s = scenes.ball2d_rsa(n=20000, seed=7)
    s["v"] = np.random.default_rng(7).normal(0.0, 2.0, s["v"].size)
    return s


@pytest.fixture(scope="module")
def pile():
    return scenes.spheres_pile(n=8000, seed=9)


def _run(monkeypatch, s, skip, blocks, warm, steps=5):
    monkeypatch.setenv("SCX_VIOLATOR_SKIP", "2" if skip else "0")
    if blocks:
        monkeypatch.setenv("SCX_FORCE_PERSISTENT_BLOCKS", str(blocks))
    else:
        monkeypatch.delenv("SCX_FORCE_PERSISTENT_BLOCKS", raising=False)
    D = s["dim"]
    ctx = scx.Context(D, s["r"].size, max_planes=8)
    try:
        ctx.set_autotune(False)
        ctx.set_planes(s["px"], s["pn"])
        ctx.set_bodies(s["q"], s["v"], s["r"], s["m"], inertia=s.get("inertia"))
        g = [0.0, -9.81] if D == 2 else [0.0, -9.81, 0.0]
        # cold steps at a tight tolerance: long FP32 phases that end at their rounding floor
        # through the stagnation test, then the FP64 polish
        tol = 1e-6 if warm != scx.WARM_NONE else 1e-10
        prm = dict(dt=0.005, g=g, mu=0.4, CoR=0.0, tol=tol, max_iters=20000, eval_every=1,
                   integrator=0 if D == 2 else 1, warm_start=warm, precision=scx.FP32_POLISH)
        out = []
        for k in range(steps):
            st = ctx.step(scx.step_params(iteration=k + 1, **prm))
            q, v, w = ctx.get_state()
            con = ctx.contacts()
            out.append((st, ctx.sweep_counters(), q, v, w, con["alpha"], con["beta"]))
        return out
    finally:
        ctx.close()


@pytest.mark.parametrize("warm", [scx.WARM_TRANSPORT, scx.WARM_NONE], ids=["transport", "cold"])
@pytest.mark.parametrize("blocks", [16, 0])
@pytest.mark.parametrize("kind", ["circles", "pile"])
def test_skipped_evaluations_change_nothing(monkeypatch, request, kind, blocks, warm):
    s = request.getfixturevalue(kind)
    _say("case", kind, "blocks", blocks, "warm", warm)
    steps = 5 if warm != scx.WARM_NONE else 3
    ref = _run(monkeypatch, s, False, blocks, warm, steps)
    got = _run(monkeypatch, s, True, blocks, warm, steps)
    long_fp32 = False
    for k, (a, b) in enumerate(zip(ref, got)):
        _say("step", k, "contacts", a[0]["ncontacts"], "colours", a[0]["ncolours"], "sweeps", a[0]["sweeps"],
             b[0]["sweeps"], "fp32", a[0]["sweeps_fp32"], b[0]["sweeps_fp32"], "error", a[0]["error"], b[0]["error"])
        for key in ("ncontacts", "ncolours", "sweeps", "sweeps_fp32", "error", "succeeded"):
            assert a[0][key] == b[0][key], key
        for x, y in zip(a[2:], b[2:]):
            if x is not None:
                assert np.array_equal(x, y)
        st, ev_ref, ev = a[0], a[1], b[1]
        _say("evaluations", ev_ref, ev)
        if st["ncontacts"] > 0:
            s64 = st["sweeps"] - st["sweeps_fp32"]
            assert ev_ref["skipped"] == 0
            assert ev_ref["full_fp32"] == 1 + st["sweeps_fp32"]
            assert ev_ref["full"] - ev_ref["full_fp32"] == 1 + s64
            assert ev["full_fp32"] + ev["skipped"] == ev_ref["full_fp32"]
            assert ev["full"] - ev["full_fp32"] == 1 + s64
        long_fp32 = long_fp32 or a[0]["sweeps_fp32"] > 1024
    if warm == scx.WARM_NONE:
        assert long_fp32  # a stagnation test ran

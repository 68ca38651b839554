"""The sweeps' shortcut past full residual evaluations.

A full evaluation also lists the contacts whose residual is within a factor 10 of its
maximum (at least the stopping value). While workgroup 0 finds one of them at or above a
threshold after a sweep, the maximum is too, and that sweep's full evaluation is skipped:
with the threshold at the best residual so far it could only have said "go on, no new
best", so neither the decision to stop nor the best iterate (returned at the sweep limit)
nor the best residual changes. In the first 768 sweeps of the FP32 phase, which keeps no
best iterate, the threshold is the stopping value; the FP32 phase's stagnation test (every
1024 sweeps) always gets a full evaluation. The iterates, the sweep counts and the
residuals must therefore be those of a full evaluation after every sweep
(SCX_VIOLATOR_SKIP=0) bit for bit: warm-started FP32 + FP64-polish steps, cold ones whose
FP32 phases end in the stagnation test, FP64 steps, and FP64 steps at a sweep limit that
return the best iterate. SCX_VIOLATOR_SKIP=2 fails a step whose FP32 phase ran 64 sweeps or
more without skipping one. The evaluation counters (scx_get_sweep_counters) must add up:
one evaluation before the first sweep of each phase and one after every sweep
(eval_every = 1) without skipping, and with skipping the same number of full and skipped
ones together.
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


def _run(monkeypatch, s, skip, blocks, mode, steps=5):
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
        # through the stagnation test, then the FP64 polish; FP64 steps to the tolerance and
        # to a sweep limit (best iterate)
        warm = scx.WARM_NONE if mode == "cold" else scx.WARM_TRANSPORT
        tol = {"transport": 1e-6, "cold": 1e-10, "fp64": 1e-7, "fp64_limit": 1e-14}[mode]
        prec = scx.FP32_POLISH if mode in ("transport", "cold") else scx.FP64
        prm = dict(dt=0.005, g=g, mu=0.4, CoR=0.0, tol=tol, max_iters=300 if mode == "fp64_limit" else 20000,
                   eval_every=1, integrator=0 if D == 2 else 1, warm_start=warm, precision=prec)
        out = []
        for k in range(steps):
            st = ctx.step(scx.step_params(iteration=k + 1, **prm))
            q, v, w = ctx.get_state()
            con = ctx.contacts()
            out.append((st, ctx.sweep_counters(), q, v, w, con["alpha"], con["beta"]))
        return out
    finally:
        ctx.close()


@pytest.mark.parametrize("mode", ["transport", "cold", "fp64", "fp64_limit"])
@pytest.mark.parametrize("blocks", [16, 0])
@pytest.mark.parametrize("kind", ["circles", "pile"])
def test_skipped_evaluations_change_nothing(monkeypatch, request, kind, blocks, mode):
    s = request.getfixturevalue(kind)
    _say("case", kind, "blocks", blocks, "mode", mode)
    steps = 3 if mode == "cold" else 5
    ref = _run(monkeypatch, s, False, blocks, mode, steps)
    got = _run(monkeypatch, s, True, blocks, mode, steps)
    long_fp32 = False
    limited = skipped = 0
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
            if mode in ("transport", "cold"):
                assert ev_ref["full_fp32"] == 1 + st["sweeps_fp32"]
                assert ev_ref["full"] - ev_ref["full_fp32"] == 1 + s64
            else:
                assert ev_ref["full"] == 1 + st["sweeps"]
            assert ev["full"] + ev["skipped"] == ev_ref["full"]
        long_fp32 = long_fp32 or a[0]["sweeps_fp32"] > 1024
        limited += int(not st["succeeded"])
        skipped += b[1]["skipped"]
    _say("skipped evaluations", skipped)
    if mode == "cold":
        assert long_fp32  # a stagnation test ran
    if mode == "fp64_limit":
        assert limited >= 2  # best iterates returned

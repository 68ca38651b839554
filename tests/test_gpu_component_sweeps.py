"""The component phase of the persistent sweep kernel.

For one scene swept with coherent passes on several workgroups, the late colours (those
after which at most 20% of the contacts remain) are grouped at the start of every sweep
launch into connected components through their bodies (a union-find over the bodies), and
every component goes to one workgroup, which sweeps its contacts colour by colour through
its L2 between an agent-scope acquire and release, with one grid barrier for the whole
phase. Components share no body and every contact is still updated in its colour's turn,
so the iterates must be those of the grid-wide passes and the one-workgroup tail
(SCX_COMPONENT_SWEEPS=0) bit for bit: positions, velocities, impulses, sweep counts and
residuals, on a 2D circle packing and a 3D sphere pile, in FP64 and in FP32 sweeps with
the FP64 polish (whose skipped evaluations then follow a check after the phase), on 16 and
64 workgroups and on the width the cost model picks, with steps that converge and with a
sweep limit that returns the best iterate; with each workgroup's components held in LDS for
the launch (bodies loaded and stored coherently around the phase, impulses kept in LDS and
written through) and, with SCX_COMPONENT_LDS=0, swept through the L2 with ordinary accesses
between an acquire and a release. The sweep counters must show the phase in the
steps of the one run and in none of the other, so that the comparison is not between two
runs of the same path. The circle packings have five or six colours, whose last 20% of the
contacts often lie in one colour (then there is no phase): they run with the share raised to
35% (SCX_COMPONENT_PERCENT), the pile with the default.
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
    s = scenes.ball2d_rsa(n=20000, seed=11)
    s["v"] = np.random.default_rng(11).normal(0.0, 2.0, s["v"].size)  # collisions from the first step
    return s


@pytest.fixture(scope="module")
def pile():
    return scenes.spheres_pile(n=8000, seed=13)


def _run(monkeypatch, kind, s, mode, blocks, precision, steps, limit, lds=True):
    monkeypatch.setenv("SCX_COMPONENT_SWEEPS", str(mode))
    monkeypatch.setenv("SCX_COMPONENT_LDS", "1" if lds else "0")
    monkeypatch.setenv("SCX_COMPONENT_PERCENT", "35" if kind == "circles" else "20")
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
        prm = dict(dt=0.005, g=g, mu=0.4, CoR=0.0, eval_every=1, integrator=0 if D == 2 else 1,
                   warm_start=scx.WARM_TRANSPORT, precision=precision)
        prm.update(dict(tol=1e-7, max_iters=4000) if limit == "converged" else dict(tol=1e-13, max_iters=150))
        out = []
        for k in range(steps):
            st = ctx.step(scx.step_params(iteration=k + 1, **prm))
            q, v, w = ctx.get_state()
            con = ctx.contacts()
            out.append((st, ctx.sweep_counters(), q, v, w, con["alpha"], con["beta"]))
        return out
    finally:
        ctx.close()


@pytest.mark.parametrize("where", ["lds", "l2"])
@pytest.mark.parametrize("limit", ["converged", "sweep_limit"])
@pytest.mark.parametrize("precision", [scx.FP32_POLISH, scx.FP64], ids=["fp32_polish", "fp64"])
@pytest.mark.parametrize("blocks", [16, 64, 0])
@pytest.mark.parametrize("kind", ["circles", "pile"])
def test_component_phase_matches_colour_passes_bit_for_bit(monkeypatch, request, kind, blocks, precision, limit,
                                                           where):
    s = request.getfixturevalue(kind)
    _say("case", kind, "blocks", blocks, "precision", precision, "limit", limit, "where", where)
    ref = _run(monkeypatch, kind, s, 0, blocks, precision, 4, limit)
    got = _run(monkeypatch, kind, s, 1, blocks, precision, 4, limit, lds=where == "lds")
    swept = limited = with_phase = 0
    for k, (a, b) in enumerate(zip(ref, got)):
        _say("step", k, "contacts", a[0]["ncontacts"], "colours", a[0]["ncolours"], "sweeps", a[0]["sweeps"],
             b[0]["sweeps"], "error", a[0]["error"], b[0]["error"], "component launches",
             a[1]["component_launches"], b[1]["component_launches"])
        for key in ("ncontacts", "ncolours", "sweeps", "sweeps_fp32", "error", "succeeded"):
            assert a[0][key] == b[0][key], key
        for x, y in zip(a[2:], b[2:]):
            if x is not None:
                assert np.array_equal(x, y)
        assert a[1]["component_launches"] == 0
        swept += int(a[0]["sweeps"] > 2)
        limited += int(not a[0]["succeeded"])
        with_phase += int(b[1]["component_launches"] > 0 and a[0]["sweeps"] > 2)
    assert swept >= 2
    assert with_phase >= 2
    if limit == "sweep_limit":
        assert limited >= 2

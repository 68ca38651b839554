"""The colour passes with coherent body accesses and light grid barriers.

Between two colour passes the persistent sweep kernel reads and writes the body records
with coherent (SC1) buffer accesses and synchronises with grid_sync_coherent, which skips
the L2 writeback and invalidate of the full grid barrier; evaluations, the best-iterate
return and compactions still follow full barriers. The arithmetic is the same, so the
iterates must be those of the full-barrier kernel (SCX_COHERENT_SWEEPS=0) bit for bit:
positions, velocities, impulses, sweep counts and residuals, over several steps of a 2D
circle packing, a 3D sphere pile and an ensemble, in FP64 and in FP32 sweeps with the FP64
polish, on one workgroup, on 16 and 64 workgroups and on the width the cost model picks, with
steps that converge and with a sweep limit that returns the best iterate.
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
    s = scenes.ball2d_rsa(n=20000, seed=5)
    s["v"] = np.random.default_rng(5).normal(0.0, 2.0, s["v"].size)  # collisions from the first step
    return s


@pytest.fixture(scope="module")
def pile():
    return scenes.spheres_pile(n=8000, seed=9)


@pytest.fixture(scope="module")
def ensemble():
    return scenes.ensemble_piles(nscenes=32, nmin=200, nmax=2000, seed=4, workers=8)


def _context(kind, data):
    if kind == "circles":
        s = data
        ctx = scx.Context(2, s["r"].size, max_planes=8)
        ctx.set_planes(s["px"], s["pn"])
        ctx.set_bodies(s["q"], s["v"], s["r"], s["m"])
        g = [0.0, -9.81]
    elif kind == "pile":
        s = data
        ctx = scx.Context(3, s["r"].size, max_planes=8)
        ctx.set_planes(s["px"], s["pn"])
        ctx.set_bodies(s["q"], s["v"], s["r"], s["m"], inertia=s["inertia"])
        g = [0.0, -9.81, 0.0]
    else:
        cat, bo, po, mus, _ = data
        ctx = scx.Context(3, int(bo[-1]), max_planes=int(po[-1]))
        ctx.set_planes(cat["px"], cat["pn"])
        ctx.set_bodies(cat["q"], cat["v"], cat["r"], cat["m"], inertia=cat["inertia"])
        ctx.set_scenes(bo, po, mus)
        g = [0.0, -9.81, 0.0]
    ctx.set_autotune(False)
    return ctx, g


def _run(monkeypatch, kind, data, coherent, blocks, precision, steps, limit):
    monkeypatch.setenv("SCX_COHERENT_SWEEPS", "1" if coherent else "0")
    if blocks:
        monkeypatch.setenv("SCX_FORCE_PERSISTENT_BLOCKS", str(blocks))
    else:
        monkeypatch.delenv("SCX_FORCE_PERSISTENT_BLOCKS", raising=False)
    ctx, g = _context(kind, data)
    try:
        prm = dict(dt=0.005, g=g, mu=0.4, CoR=0.0, eval_every=1, integrator=0 if kind == "circles" else 1,
                   warm_start=scx.WARM_TRANSPORT, precision=precision)
        # converged steps, or a sweep limit that returns the best iterate
        prm.update(dict(tol=1e-7, max_iters=4000) if limit == "converged" else dict(tol=1e-13, max_iters=150))
        out = []
        for k in range(steps):
            st = ctx.step(scx.step_params(iteration=k + 1, **prm))
            q, v, w = ctx.get_state()
            con = ctx.contacts()
            out.append((st, q, v, w, con["alpha"], con["beta"]))
        return out
    finally:
        ctx.close()


@pytest.mark.parametrize("limit", ["converged", "sweep_limit"])
@pytest.mark.parametrize("precision", [scx.FP32_POLISH, scx.FP64], ids=["fp32_polish", "fp64"])
@pytest.mark.parametrize("blocks", [1, 16, 64, 0])
@pytest.mark.parametrize("kind", ["circles", "pile", "ensemble"])
def test_coherent_passes_match_full_barriers_bit_for_bit(monkeypatch, request, kind, blocks, precision, limit):
    data = request.getfixturevalue(kind)
    _say("case", kind, "blocks", blocks, "precision", precision, "limit", limit)
    ref = _run(monkeypatch, kind, data, False, blocks, precision, 4, limit)
    got = _run(monkeypatch, kind, data, True, blocks, precision, 4, limit)
    swept = limited = 0
    for k, (a, b) in enumerate(zip(ref, got)):
        _say("step", k, "contacts", a[0]["ncontacts"], "colours", a[0]["ncolours"], "sweeps", a[0]["sweeps"],
             b[0]["sweeps"], "error", a[0]["error"], b[0]["error"])
        for key in ("ncontacts", "ncolours", "sweeps", "sweeps_fp32", "error", "succeeded"):
            assert a[0][key] == b[0][key], key
        for x, y in zip(a[1:], b[1:]):
            if x is not None:
                assert np.array_equal(x, y)
        swept += int(a[0]["sweeps"] > 2)
        limited += int(not a[0]["succeeded"])
    assert swept >= 2
    if limit == "sweep_limit":
        assert limited >= 2

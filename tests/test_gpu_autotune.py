"""Online autotuning of the persistent sweep launches (scx_set_autotune).

The tuner only changes how many workgroups a sweep launch runs on; the contacts of one
colour are independent and the residual is a maximum, so every step must be bit-identical
with the tuner on and off while it explores several widths.
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


def _params(**kw):
    d = dict(dt=0.005, g=[0.0, -9.81, 0.0], mu=0.4, CoR=0.0, tol=1e-9, max_iters=4000, eval_every=1, integrator=1,
             precision=scx.FP32_POLISH, warm_start=scx.WARM_TRANSPORT)
    d.update(kw)
    return d


def _state(ctx):
    q, v, w = ctx.get_state()
    con = ctx.contacts()
    return q, v, w, con["alpha"], con["beta"]


def _run_pair(make, steps, prm):
    """The same scene stepped by a tuned and an untuned context; per step the states and
    impulses must be identical. Returns the tuned context's table after each step."""
    tuned, plain = make(), make()
    tuned.set_autotune(True)
    tables, widths = [], set()
    for k in range(steps):
        st_t = tuned.step(scx.step_params(**prm))
        st_p = plain.step(scx.step_params(**prm))
        tab = tuned.autotune_table()
        tables.append(tab)
        for e in tab:
            widths.add((e["key"], e["blocks"]))
        _say("step", k, "contacts", st_t["ncontacts"], "sweeps", st_t["sweeps"], st_p["sweeps"], "table",
             [(e["phase"], e["contacts_half_octave"], e["exponent"], e["blocks"], round(e["us_per_sweep"], 1),
               e["settled"]) for e in tab])
        assert st_t["sweeps"] == st_p["sweeps"] and st_t["sweeps_fp32"] == st_p["sweeps_fp32"]
        assert st_t["error"] == st_p["error"]
        for a, b in zip(_state(tuned), _state(plain)):
            assert np.array_equal(a, b)
    tuned.close()
    plain.close()
    return tables, widths


def test_tuned_and_untuned_steps_are_identical():
    s = scenes.spheres_pile(n=3000, rmed=0.05, sigma=0.25, seed=7)
    n = s["r"].size

    def make():
        c = scx.Context(3, n, max_planes=8)
        c.set_planes(s["px"], s["pn"])
        c.set_bodies(s["q"], s["v"], s["r"], s["m"], inertia=s["inertia"])
        return c

    tables, widths = _run_pair(make, 10, _params())
    last = tables[-1]
    assert last, "no sweep launch went through the tuner"
    timed = [e for e in last if e["us_per_sweep"] > 0]
    assert timed, "no launch was timed"
    # the tuner tried more than one width for some class while the results stayed identical
    keys = {k for k, _ in widths}
    assert any(len({b for kk, b in widths if kk == k}) > 1 for k in keys), widths


def test_tuned_and_untuned_2d_steps_are_identical():
    """Circles on a nearly touching lattice (many contacts from the first steps, 2D records
    and a different tail of colours)."""
    s = scenes.ball2d_lattice(m=100, side=10.1, r=0.05, vstd=0.1, seed=3)
    n = s["r"].size

    def make():
        c = scx.Context(2, n, max_planes=8)
        c.set_planes(s["px"], s["pn"])
        c.set_bodies(s["q"], s["v"], s["r"], s["m"])
        return c

    prm = _params(g=[0.0, -9.81], mu=0.3, integrator=0, warm_start=scx.WARM_NONE)
    tables, widths = _run_pair(make, 10, prm)
    keys = {k for k, _ in widths}
    assert any(len({b for kk, b in widths if kk == k}) > 1 for k in keys), widths


def test_tuned_ensemble_steps_are_identical():
    """Ensembles: per-scene freezing and residuals with the workgroups the tuner picks."""
    s = scenes.spheres_pile(n=1500, rmed=0.05, sigma=0.25, seed=8)
    n = s["r"].size
    q = np.concatenate([s["q"], s["q"]])
    v = np.concatenate([s["v"], s["v"]])
    r = np.concatenate([s["r"], s["r"]])
    m = np.concatenate([s["m"], s["m"]])
    inertia = np.concatenate([s["inertia"], s["inertia"]])
    px = np.concatenate([np.asarray(s["px"]).reshape(-1, 3)] * 2)
    pn = np.concatenate([np.asarray(s["pn"]).reshape(-1, 3)] * 2)
    npl = px.shape[0] // 2

    def make():
        c = scx.Context(3, 2 * n, max_planes=2 * npl)
        c.set_planes(px, pn)
        c.set_bodies(q, v, r, m, inertia=inertia)
        c.set_scenes([0, n, 2 * n], [0, npl, 2 * npl], [0.2, 0.7])
        return c

    tables, _ = _run_pair(make, 6, _params())
    assert any(e["ensemble"] for e in tables[-1])


def test_autotune_can_be_switched_off():
    s = scenes.spheres_pile(n=800, rmed=0.05, sigma=0.25, seed=9)
    c = scx.Context(3, s["r"].size, max_planes=8)
    c.set_planes(s["px"], s["pn"])
    c.set_bodies(s["q"], s["v"], s["r"], s["m"], inertia=s["inertia"])
    c.set_autotune(True)
    c.set_autotune(False)
    c.step(scx.step_params(**_params()))
    assert c.autotune_table() == []
    c.close()

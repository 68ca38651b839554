"""Ensembles (scx_set_scenes): many independent scenes stepped together.

Within each scene the active set, its order and the contact data must be those of the
scene stepped alone (bit-exact), bodies of different scenes never interact (also when
the scenes overlap in space), and the post-step velocities agree with the scene alone
(the FP64 oracle in the reference's natural order, or the library on the scene alone)
to a bound set by the residual tolerance, stated in each test.
"""

import sys

import numpy as np
import pytest

import oracle
import scisim_amd as scx
from scisim_amd import scenes

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _params(dim, **kw):
    d = dict(dt=0.01, g=[0.0, -9.81] if dim == 2 else [0.0, -9.81, 0.0], mu=0.5, CoR=0.0, tol=1e-12,
             max_iters=20000, eval_every=1, integrator=0 if dim == 2 else 1)
    d.update(kw)
    return d


def _concat(parts):
    """Scene dicts -> concatenated arrays and the body / plane offsets."""
    D = parts[0]["dim"]
    cat = {k: np.concatenate([p[k] for p in parts]) for k in ("q", "v", "r", "m")}
    if D == 3:
        cat["inertia"] = np.concatenate([p["inertia"] for p in parts])
        cat["w"] = np.concatenate([p.get("w", np.zeros(3 * p["r"].size)) for p in parts])
    cat["px"] = np.concatenate([np.asarray(p["px"]).reshape(-1, D) for p in parts])
    cat["pn"] = np.concatenate([np.asarray(p["pn"]).reshape(-1, D) for p in parts])
    bo = np.cumsum([0] + [p["r"].size for p in parts]).astype(np.uint32)
    po = np.cumsum([0] + [np.asarray(p["px"]).reshape(-1, D).shape[0] for p in parts]).astype(np.uint32)
    return cat, bo, po


def _scene_contacts(con, b0, b1, p0):
    """The contacts of one scene from an ensemble's list, in local ids."""
    sel = (con["a"] >= b0) & (con["a"] < b1)
    a = con["a"][sel] - b0
    b = np.where(con["b"][sel] >= 0, con["b"][sel] - b0, -1)
    pl = np.where(con["plane"][sel] >= 0, con["plane"][sel] - p0, -1)
    return a, b, pl, con["basis"][sel]


def _settled_lattices(sizes, mus, steps=20):
    """Small ball2d lattices run a few steps on the oracle (a loose pile, unique solutions)."""
    parts = []
    for k, (m, mu) in enumerate(zip(sizes, mus)):
        s = scenes.ball2d_lattice(m=m, side=m * 0.156, r=0.05, vstd=0.1, seed=11 + k)
        sim = oracle.Sim(2, s["q"], s["v"], s["r"], s["m"], s["px"], s["pn"])
        sim.run(steps, **_params(2, mu=mu, tol=1e-14))
        parts.append({"dim": 2, "q": sim.q.copy(), "v": sim.v.copy(), "r": s["r"], "m": s["m"], "px": s["px"],
                      "pn": s["pn"], "mu": mu})
    return parts


def test_ensemble_2d_matches_each_scene_alone():
    """Four lattices of different sizes and friction coefficients. Active sets per scene
    bit-exact against the library on the scene alone (same order, bases, planes);
    post-step velocities against the FP64 oracle on the scene alone in the natural order:
    ||FB||^2 < 1e-14 on both sides, velocities O(1), so a relative bound of 1e-5."""
    sizes, mus = [12, 20, 28, 16], [0.1, 0.3, 0.6, 0.8]
    parts = _settled_lattices(sizes, mus)
    cat, bo, po = _concat(parts)
    ens = scx.Context(2, int(bo[-1]), max_planes=int(po[-1]))
    ens.set_planes(cat["px"], cat["pn"])
    ens.set_bodies(cat["q"], cat["v"], cat["r"], cat["m"])
    ens.set_scenes(bo, po, mus)
    st = ens.step(scx.step_params(**_params(2, mu=99.0, tol=1e-14)))  # mu of the step is unused
    _say("ensemble", st)
    assert st["succeeded"]
    con = ens.contacts()
    q_e, v_e, _ = ens.get_state()
    cnt, err, sw = ens.scene_stats()
    _say("scene contacts", cnt, "residuals", err, "sweeps", sw)
    assert np.all(err < 1e-14)
    # each scene stops at its own sweep (it is frozen once it meets the tolerance)
    assert np.all(sw <= st["sweeps"]) and sw.max() == st["sweeps"]
    solo = scx.Context(2, int(max(p["r"].size for p in parts)), max_planes=8)
    for k, p in enumerate(parts):
        b0, b1 = int(bo[k]), int(bo[k + 1])
        solo.set_planes(p["px"], p["pn"])
        solo.set_bodies(p["q"], p["v"], p["r"], p["m"])
        so = solo.step(scx.step_params(**_params(2, mu=p["mu"], tol=1e-14)))
        sc = solo.contacts()
        a, b, pl, basis = _scene_contacts(con, b0, b1, int(po[k]))
        assert cnt[k] == so["ncontacts"] == a.size
        assert np.array_equal(a, sc["a"]) and np.array_equal(b, sc["b"]) and np.array_equal(pl, sc["plane"])
        assert np.array_equal(basis, sc["basis"])
        sim = oracle.Sim(2, p["q"], p["v"], p["r"], p["m"], p["px"], p["pn"])
        sim.step(order=0, **_params(2, mu=p["mu"], tol=1e-14))
        dv = np.abs(v_e[2 * b0:2 * b1] - sim.v).max() / np.abs(sim.v).max()
        _say("scene", k, "contacts", a.size, "rel v err vs oracle", dv)
        assert dv < 1e-5
        # positions q0 + dt v1: the velocity bound times dt
        assert np.abs(q_e[2 * b0:2 * b1] - sim.q).max() <= 0.01 * 1e-5 * np.abs(sim.v).max() + 1e-15
    solo.close()
    ens.close()


def test_ensemble_scenes_in_the_same_place_do_not_interact():
    """Two copies of one 3D pile at the same coordinates, with different friction: each
    copy's active set is the pile's own (bit-exact against the pile alone), none across
    copies, and each copy's velocities match the pile alone with its mu (library, same
    colouring rule; 1e-6 relative at ||FB||^2 < 1e-12)."""
    s = scenes.spheres_pile(n=400, rmed=0.05, sigma=0.25, seed=4)
    parts = []
    for mu in (0.2, 0.7):
        parts.append(dict(s, mu=mu))
    cat, bo, po = _concat(parts)
    ens = scx.Context(3, int(bo[-1]), max_planes=int(po[-1]))
    ens.set_planes(cat["px"], cat["pn"])
    ens.set_bodies(cat["q"], cat["v"], cat["r"], cat["m"], inertia=cat["inertia"])
    ens.set_scenes(bo, po, [0.2, 0.7])
    prm = _params(3, dt=0.005, tol=1e-12)
    st = ens.step(scx.step_params(**prm))
    _say("ensemble", st)
    assert st["succeeded"]
    con = ens.contacts()
    _, v_e, w_e = ens.get_state()
    solo = scx.Context(3, s["r"].size, max_planes=8)
    solo.set_planes(s["px"], s["pn"])
    n = s["r"].size
    for k, mu in enumerate((0.2, 0.7)):
        b0, b1 = int(bo[k]), int(bo[k + 1])
        a, b, pl, basis = _scene_contacts(con, b0, b1, int(po[k]))
        assert np.all((b < 0) | ((b >= 0) & (b < n)))
        solo.set_bodies(s["q"], s["v"], s["r"], s["m"], inertia=s["inertia"])
        so = solo.step(scx.step_params(**dict(prm, mu=mu)))
        sc = solo.contacts()
        assert np.array_equal(a, sc["a"]) and np.array_equal(b, sc["b"]) and np.array_equal(pl, sc["plane"])
        assert np.array_equal(basis, sc["basis"])
        _, v_s, w_s = solo.get_state()
        dv = np.abs(v_e[3 * b0:3 * b1] - v_s).max() / max(np.abs(v_s).max(), 1e-300)
        _say("copy", k, "mu", mu, "contacts", a.size, so["ncontacts"], "rel v err", dv)
        assert dv < 1e-6
    assert con["a"].size == 2 * so["ncontacts"]
    solo.close()
    ens.close()


def test_ensemble_arguments_are_checked():
    s = scenes.ball2d_lattice(m=8, side=1.25, r=0.05, vstd=0.1, seed=1)
    c = scx.Context(2, 128, max_planes=8)
    c.set_planes(s["px"], s["pn"])
    c.set_bodies(s["q"], s["v"], s["r"], s["m"])
    n = s["r"].size
    with pytest.raises(scx.ScxError):
        c.set_scenes([0, n - 1], [0, 4], [0.5])  # bodies do not add up
    with pytest.raises(scx.ScxError):
        c.set_scenes([0, n], [0, 3], [0.5])  # planes do not add up
    with pytest.raises(scx.ScxError):
        c.set_scenes([0, n], [0, 4], [-0.5])  # negative mu
    c.set_scenes([0, 10, n], [0, 0, 4], [0.5, 0.1])
    st = c.step(scx.step_params(**_params(2, tol=1e-10)))
    cnt, err, sw = c.scene_stats()
    assert cnt.sum() == st["ncontacts"] and np.all(err < 1e-10) and np.all(sw <= st["sweeps"])
    c.set_scenes([], [], [])  # back to one scene
    with pytest.raises(scx.ScxError):
        c.scene_stats()
    c.close()

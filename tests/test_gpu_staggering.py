"""GPU parity of the symplectic-Euler and stabilized staggerings
(scisim/ConstrainedMaps/SymplecticEulerImpactFrictionMap.cpp:201-407,
StabilizedImpactFrictionMap.cpp:48-183) against the FP64 oracle, whose restatement of
them is pinned by closed forms in tests/test_oracle_staggering.py.

The device folds the symplectic-Euler force increment into the solve's start velocity
(v0 + dt M^-1 F) instead of carrying nrel = N^T vdelta, drel = D^T vdelta as forcing
terms: the same problem in exact arithmetic, so in FP64 and the oracle's colour order the
active set, order, bases and colours are bit-exact and impulses, velocities and
positions agree at rounding level (relative 1e-9) whenever both solves stop within a
sweep of each other; a slowly converging solve whose rounding differences move its
crossing of the tolerance is judged at the tolerance (see _lockstep).
"""

import sys

import numpy as np
import pytest

import oracle
import scisim_amd as scx
from parity import check_with_oracle, position_excess, rel, same_active_set
from scisim_amd import scenes

pytestmark = pytest.mark.gpu

G = 9.81
P2 = dict(dt=0.01, g=[0.0, -G], mu=0.5, CoR=0.0, tol=1e-10, max_iters=20000, eval_every=1, integrator=0)
P3 = dict(dt=0.005, g=[0.0, -G, 0.0], mu=0.4, CoR=0.0, tol=1e-10, max_iters=20000, eval_every=1, integrator=1)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _solve_start(sim, prm, stag_kw):
    """The velocity the step's solve starts from in the device's formulation: v0 + dt M^-1 F
    (symplectic Euler, force increment folded in), the unconstrained predictor velocity
    (stabilized), else v0."""
    D = sim.dim
    stag = stag_kw.get("staggering", 0)
    g = np.array(prm["g"][:D])
    if stag == scx.STAGGER_SYMPLECTIC_EULER:
        m = sim.m
        return (sim.v.reshape(-1, D) + prm["dt"] * ((1.0 / m)[:, None] * (m[:, None] * g))).ravel()
    if stag == scx.STAGGER_STABILIZED:
        return oracle.umap_flow(D, prm["integrator"], sim.m, g, prm["dt"], sim.q, sim.v)[1]
    return sim.v.copy()


def _lockstep(ctx, sim, prm, warm, nsteps, tag, stag_kw):
    """nsteps FP64 steps of both from the oracle's state (oracle in colour order); returns
    the worst relative errors of the steps that stopped within one sweep of each other.
    Where a slowly converging solve crosses the tolerance at sweeps further apart (rounding
    differences steer the iterates), the step is judged at the tolerance instead: both meet
    it, the oracle's FB residual of the device's impulses meets it, the device's velocities
    are those impulses' (rounding level), and the two results lie within the distance the
    tolerance allows (the larger of 4 sqrt(tol) and ten times the distance between the
    oracle's own solutions in its natural and colour orders, in velocity)."""
    worst = {}
    for k in range(nsteps):
        state0 = dict(q=sim.q.copy(), v=_solve_start(sim, prm, stag_kw), w=None if sim.w is None else sim.w.copy(),
                      r=sim.r, m=sim.m, inertia=sim.inertia)
        alt = sim.clone()  # the same state and impulse cache, for the oracle's natural order
        ctx.set_bodies(sim.q, sim.v, sim.r, sim.m, omega=sim.w, inertia=sim.inertia)
        st = ctx.step(scx.step_params(precision=scx.FP64, warm_start=warm, **prm, **stag_kw))
        so = sim.step(order=1, warm_start=warm, **prm, **stag_kw)
        so["sweeps"] = min(so["sweeps"], prm["max_iters"])  # So-bogus's GSIter ends one past the limit
        gc, oc = ctx.contacts(), sim.contacts()
        _say(tag, "step", k, "contacts", st["ncontacts"], "colours", st["ncolours"], "sweeps", st["sweeps"],
             so["sweeps"], "residual", st["error"], so["error"])
        assert st["ncontacts"] == so["ncontacts"] > 0
        assert same_active_set(gc, oc)
        assert np.array_equal(gc["basis"], oc["basis"])
        assert np.array_equal(gc["colour"], oc["colour"])
        worst["steps"] = worst.get("steps", 0) + 1
        if abs(st["sweeps"] - so["sweeps"]) > 1:
            err, vout = check_with_oracle(sim.dim, state0, gc, prm["mu"], prm.get("CoR", 0.0))
            v2 = ctx.v2()[0]
            vd = vout.reshape(-1, 6)[:, :3].ravel() if sim.dim == 3 else vout
            dist = float(np.abs(v2 - sim.v2()).max())
            # two solutions that both meet the tolerance differ by as much as the oracle's own
            # solutions in its two orders (natural and colour) may
            alt.step(order=0, warm_start=warm, **prm, **stag_kw)
            spread = float(np.abs(alt.v2() - sim.v2()).max())
            bound = max(4.0 * np.sqrt(prm["tol"]), 10.0 * spread)
            _say(tag, "step", k, "judged at the tolerance: oracle residual of the device impulses", err,
                 "|v2 - oracle v2| %.3e, oracle order spread %.3e, bound %.3e" % (dist, spread, bound))
            assert st["succeeded"] and so["succeeded"] and err < prm["tol"] * 1.0001
            assert rel(v2, vd) < 1e-9
            assert dist <= bound
            worst["steps_at_tolerance"] = worst.get("steps_at_tolerance", 0) + 1
            continue
        if st["sweeps"] != so["sweeps"]:
            continue
        q, v, w = ctx.get_state()
        amax = max(1e-300, np.abs(oc["alpha"]).max())
        e = {"v2": rel(ctx.v2()[0], sim.v2()), "v1": rel(v, sim.v),
             "q1": position_excess(q, sim.q, [(v, sim.v), (ctx.v2()[0], sim.v2())], prm["dt"])}
        # impulses are compared where both solves converged: at the sweep limit the best
        # iterates of a statically indeterminate contact set (a body resting on several
        # contacts) may differ along the impulses that leave every velocity unchanged
        imp = {"alpha": rel(gc["alpha"], oc["alpha"]), "beta": rel(gc["beta"], oc["beta"], amax)}
        if st["succeeded"] and so["succeeded"]:
            e.update(imp)
        else:
            _say(tag, "step", k, "at the sweep limit: impulse distance (reported only)", imp)
        if sim.dim == 3:
            e["omega"] = rel(w, sim.w)
        for key, x in e.items():
            worst[key] = max(worst.get(key, 0.0), x)
        worst["steps_compared"] = worst.get("steps_compared", 0) + 1
    _say(tag, "worst relative error vs the FP64 oracle:", worst)
    assert worst.get("steps_compared", 0) + worst.get("steps_at_tolerance", 0) >= max(1, int(0.6 * worst["steps"]))
    for key in ("alpha", "beta", "v2", "v1", "omega"):
        if key in worst:
            assert worst[key] < 1e-9, (key, worst[key])
    if "q1" in worst:
        assert worst["q1"] <= 1.0
    return worst


def _lattice_sim(stag_kw, prm, warm, nsteps=40, seed=1):
    s = scenes.ball2d_lattice(m=32, side=5.0, r=0.05, vstd=0.1, seed=seed)
    sim = oracle.Sim(2, s["q"], s["v"], s["r"], s["m"], s["px"], s["pn"])
    sim.run(nsteps, warm_start=warm, **prm, **stag_kw)
    return s, sim


@pytest.mark.parametrize("stag_kw,warm,integrator", [
    (dict(staggering=scx.STAGGER_SYMPLECTIC_EULER), scx.WARM_NONE, 0),
    (dict(staggering=scx.STAGGER_SYMPLECTIC_EULER), scx.WARM_CACHE, 0),
    (dict(staggering=scx.STAGGER_SYMPLECTIC_EULER, stabilize=1, penetration_threshold=1e-4), scx.WARM_NORMAL, 0),
    (dict(staggering=scx.STAGGER_STABILIZED), scx.WARM_NONE, 0),
    (dict(staggering=scx.STAGGER_STABILIZED), scx.WARM_NONE, 1),
], ids=["se", "se_cache", "se_stabilize_normal", "stabilized_se", "stabilized_verlet"])
def test_circles_fp64_stepwise(stag_kw, warm, integrator):
    prm = dict(P2, integrator=integrator)
    s, sim = _lattice_sim(stag_kw, prm, warm)
    ctx = scx.Context(2, s["r"].size, max_planes=4)
    try:
        ctx.set_planes(s["px"], s["pn"])
        # one step on each side fills both caches from the same state
        ctx.set_bodies(sim.q, sim.v, sim.r, sim.m)
        ctx.step(scx.step_params(precision=scx.FP64, warm_start=warm, **prm, **stag_kw))
        sim.step(order=1, warm_start=warm, **prm, **stag_kw)
        _lockstep(ctx, sim, prm, warm, 4, "2D %s warm %d" % (stag_kw, warm), stag_kw)
    finally:
        ctx.close()


@pytest.fixture(scope="module")
def pile2k():
    """A poured pile of 2,000 spheres, settled on the device under the symplectic Euler
    staggering with the transported cache (the reference's 3D warm-start path)."""
    s = scenes.spheres_pile(n=2000, rmed=0.05, sigma=0.25, seed=5)
    ctx = scx.Context(3, s["r"].size, max_planes=8)
    try:
        ctx.set_planes(s["px"], s["pn"])
        ctx.set_bodies(s["q"], s["v"], s["r"], s["m"], inertia=s["inertia"])
        p = scx.step_params(precision=scx.FP32_POLISH, warm_start=scx.WARM_TRANSPORT,
                            staggering=scx.STAGGER_SYMPLECTIC_EULER, **dict(P3, tol=1e-6, max_iters=2000))
        for _ in range(40):
            ctx.step(p)
        q, v, w = ctx.get_state()
    finally:
        ctx.close()
    return dict(s, q=q, v=v, w=w)


@pytest.mark.parametrize("stag_kw,warm", [
    (dict(staggering=scx.STAGGER_SYMPLECTIC_EULER), scx.WARM_TRANSPORT),
    (dict(staggering=scx.STAGGER_SYMPLECTIC_EULER, stabilize=1, penetration_threshold=1e-5), scx.WARM_NONE),
    (dict(staggering=scx.STAGGER_STABILIZED), scx.WARM_NONE),
], ids=["se_transport", "se_stabilize", "stabilized"])
def test_spheres_fp64_stepwise(pile2k, stag_kw, warm):
    s = pile2k
    sim = oracle.Sim(3, s["q"], s["v"], s["r"], s["m"], s["px"], s["pn"], w=s["w"], inertia=s["inertia"])
    ctx = scx.Context(3, s["r"].size, max_planes=8)
    try:
        ctx.set_planes(s["px"], s["pn"])
        ctx.set_bodies(sim.q, sim.v, sim.r, sim.m, omega=sim.w, inertia=sim.inertia)
        ctx.step(scx.step_params(precision=scx.FP64, warm_start=warm, **P3, **stag_kw))
        sim.step(order=1, warm_start=warm, **P3, **stag_kw)
        _lockstep(ctx, sim, P3, warm, 3, "3D %s warm %d" % (stag_kw, warm), stag_kw)
    finally:
        ctx.close()


def test_symplectic_euler_fp32_polish_meets_tolerance(pile2k):
    """The benchmark's precision (FP32 sweeps + FP64 polish) under the symplectic Euler
    staggering with the transported cache: the oracle's FP64 residual of the device's
    impulses, on the equivalent problem started from v0 + dt M^-1 F, meets 1e-6, and the
    device's v1 is that start velocity plus M^-1 H^T r at rounding level."""
    s = pile2k
    prm = dict(P3, tol=1e-6)
    ctx = scx.Context(3, s["r"].size, max_planes=8)
    try:
        ctx.set_planes(s["px"], s["pn"])
        ctx.set_bodies(s["q"], s["v"], s["r"], s["m"], omega=s["w"], inertia=s["inertia"])
        tp = scx.step_params(precision=scx.FP32_POLISH, warm_start=scx.WARM_TRANSPORT,
                             staggering=scx.STAGGER_SYMPLECTIC_EULER, **prm)
        ctx.step(tp)
        for k in range(3):
            q, v, w = ctx.get_state()
            vstart = (v.reshape(-1, 3) + prm["dt"] * ((1.0 / s["m"])[:, None] * (s["m"][:, None] *
                                                                               np.array(prm["g"])))).ravel()
            state0 = dict(q=q, v=vstart, w=w, r=s["r"], m=s["m"], inertia=s["inertia"])
            st = ctx.step(tp)
            gc = ctx.contacts()
            err, vout = check_with_oracle(3, state0, gc, prm["mu"])
            v1, w1 = ctx.v2()
            _say("SE FP32 polish step", k, "contacts", st["ncontacts"], "sweeps", st["sweeps"], "device residual",
                 st["error"], "oracle residual of the device impulses", err)
            assert st["succeeded"] and err < prm["tol"] * 1.0001
            assert rel(v1, vout.reshape(-1, 6)[:, :3]) < 1e-9
            assert rel(w1, vout.reshape(-1, 6)[:, 3:]) < 1e-9
    finally:
        ctx.close()


@pytest.mark.parametrize("theta_deg,mu", [(20, 0.5), (35, 0.1)])
def test_sphere_on_incline_closed_form_on_the_device(theta_deg, mu):
    """Rolling (mu >= 2/7 tan theta) or sliding from the first step under the symplectic
    Euler staggering, through the device."""
    theta = np.radians(theta_deg)
    n = np.array([-np.sin(theta), np.cos(theta), 0.0])
    d = np.array([-np.cos(theta), -np.sin(theta), 0.0])
    rad, m, dt = 0.1, 2.0, 0.005
    ctx = scx.Context(3, 1, max_planes=4)
    try:
        ctx.set_planes(np.zeros(3), n)
        ctx.set_bodies(rad * n, np.zeros(3), np.array([rad]), np.array([m]))
        rolling = mu >= 2.0 / 7.0 * np.tan(theta)
        a = 5.0 / 7.0 * G * np.sin(theta) if rolling else G * (np.sin(theta) - mu * np.cos(theta))
        for k in range(1, 11):
            st = ctx.step(scx.step_params(dt=dt, g=[0.0, -G, 0.0], mu=mu, tol=1e-26, max_iters=100, integrator=1,
                                          precision=scx.FP64, staggering=scx.STAGGER_SYMPLECTIC_EULER))
            assert st["ncontacts"] == 1
            q, v, w = ctx.get_state()
            assert abs(v @ d - k * a * dt) < 1e-12 * max(1.0, a)
            assert abs(v @ n) < 1e-12
    finally:
        ctx.close()


def test_stabilized_empty_active_set_keeps_the_unconstrained_predictor():
    """With no contact the stabilized map returns the unconstrained map's step
    (StabilizedImpactFrictionMap.cpp:66-78): Verlet under constant gravity is the exact
    parabola; symplectic Euler staggering gives q0 + dt (v0 + g dt)."""
    rng = np.random.default_rng(2)
    n = 64
    q0 = rng.uniform(-50, 50, (n, 3))
    v0 = rng.normal(0, 1, (n, 3))
    r = np.full(n, 0.01)
    m = rng.uniform(0.5, 2, n)
    g = np.array([0.0, -G, 0.0])
    dt = 0.01
    for stag in (scx.STAGGER_STABILIZED, scx.STAGGER_SYMPLECTIC_EULER):
        ctx = scx.Context(3, n, max_planes=4)
        try:
            ctx.set_bodies(q0.ravel(), v0.ravel(), r, m)
            st = ctx.step(scx.step_params(dt=dt, g=list(g), mu=0.5, integrator=1, precision=scx.FP64,
                                          staggering=stag))
            assert st["ncontacts"] == 0
            sim = oracle.Sim(3, q0.ravel(), v0.ravel(), r, m, np.zeros((0, 3)), np.zeros((0, 3)))
            sim.step(dt=dt, g=list(g), mu=0.5, integrator=1, staggering=stag)
            q, v, _ = ctx.get_state()
            assert np.array_equal(q, sim.q) and np.array_equal(v, sim.v)
        finally:
            ctx.close()


def test_invalid_staggering_options_are_refused():
    ctx = scx.Context(3, 1, max_planes=4)
    try:
        ctx.set_bodies(np.array([0.0, 0.1, 0.0]), np.zeros(3), np.array([0.1]), np.array([1.0]))
        bad = [dict(CoR=0.5, staggering=scx.STAGGER_SYMPLECTIC_EULER),
               dict(warm_start=scx.WARM_TRANSPORT, staggering=scx.STAGGER_STABILIZED),
               dict(staggering=7), dict(staggering=scx.STAGGER_SYMPLECTIC_EULER, penetration_threshold=-1.0)]
        for kw in bad:
            with pytest.raises(scx.ScxError):
                ctx.step(scx.step_params(dt=0.01, g=[0, -G, 0], mu=0.5, **kw))
    finally:
        ctx.close()

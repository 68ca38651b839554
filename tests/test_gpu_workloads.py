"""GPU parity on the benchmark workloads, against the FP64 CPU oracle.

Per workload (the circle lattice of the reference's CPU method at its real size, the RSA
circle packing, the poured sphere pile, the learned warm start on that pile, the
ensemble of piles):

* FP64 in the oracle's colour order: active set, its order, bases and colours bit-exact;
  impulses, post-step velocities, positions and the residual value itself at rounding
  level (relative 1e-9) whenever both sides stop at the same sweep (they do, except where
  the residual lands within rounding of the tolerance: then within one sweep).
* The benchmark's FP32 sweeps + FP64 polish, the learned warm start and the ensemble
  (whose colour order differs from a scene alone): the oracle evaluates the device's
  impulses itself, in FP64, with its own assembly of H and M: their FB residual must meet
  the step's tolerance (the reference's convergence criterion), and the device's v2 must
  be v0 + M^-1 H^T r at rounding level. The distance to the oracle's own solution at the
  same tolerance is reported per quantity. Its bound follows from the residual: ||phi||^2
  <= tol fixes each contact's relative velocity only to within sqrt(tol) (1e-3 m/s at
  1e-6), so two solutions that both meet it may differ in v2 by a few sqrt(tol); the
  tests allow the larger of 4 sqrt(tol) and ten times the distance between the oracle's
  own solutions in its two orders (the reference's natural order and the colour order).
  Impulses of a resting pile are statically indeterminate, so their distance is reported
  only.
* Full sizes (1,048,576 circles, 262,144 spheres): active set, order, bases and colours
  bit-exact against the oracle at the start and after device-run steps.
"""

import sys

import numpy as np
import pytest

import oracle
import scisim_amd as scx
from parity import check_with_oracle, position_excess, rel, same_active_set
from scisim_amd import scenes

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CFG0 = dict(dt=0.01, g=[0.0, -9.81], mu=0.5, CoR=0.0, tol=1e-6, max_iters=5000, eval_every=1, integrator=0)
CFG1 = dict(dt=0.005, g=[0.0, -9.81], mu=0.3, CoR=0.0, tol=1e-6, max_iters=50000, eval_every=1, integrator=0)
CFG2 = dict(dt=0.005, g=[0.0, -9.81, 0.0], mu=0.4, CoR=0.0, tol=1e-6, max_iters=50000, eval_every=1, integrator=1)


def _state(sim):
    return dict(q=sim.q.copy(), v=sim.v.copy(), w=None if sim.w is None else sim.w.copy(), r=sim.r, m=sim.m,
                inertia=sim.inertia)


def _fp64_step(ctx, sim, prm, worst, tag, warm=0):
    """One FP64 step of both from the oracle's state, oracle in colour order."""
    ctx.set_bodies(sim.q, sim.v, sim.r, sim.m, omega=sim.w, inertia=sim.inertia)
    st = ctx.step(scx.step_params(precision=scx.FP64, warm_start=warm, **prm))
    so = sim.step(order=1, warm_start=warm, **prm)
    # the oracle reports So-bogus's GSIter, one past the limit when the limit is reached
    # (scripts_include/sobogus_patch.patch:67-70); the device reports the sweeps it ran
    so["sweeps"] = min(so["sweeps"], prm["max_iters"])
    gc, oc = ctx.contacts(), sim.contacts()
    _say(tag, "contacts", st["ncontacts"], "colours", st["ncolours"], "sweeps", st["sweeps"], so["sweeps"],
         "residual", st["error"], so["error"])
    assert st["ncontacts"] == so["ncontacts"]
    assert same_active_set(gc, oc)
    assert np.array_equal(gc["basis"], oc["basis"])
    assert np.array_equal(gc["colour"], oc["colour"])
    assert abs(st["sweeps"] - so["sweeps"]) <= 1
    if st["sweeps"] == so["sweeps"] and st["ncontacts"] > 0:
        q, v, w = ctx.get_state()
        amax = max(1e-300, np.abs(oc["alpha"]).max())
        e = {"alpha": rel(gc["alpha"], oc["alpha"]), "beta": rel(gc["beta"], oc["beta"], amax),
             "v2": rel(ctx.v2()[0], sim.v2()), "v1": rel(v, sim.v),
             "q1": position_excess(q, sim.q, [(v, sim.v), (ctx.v2()[0], sim.v2())], prm["dt"]),
             "residual": abs(st["error"] - so["error"]) / max(so["error"], 1e-12 * prm["tol"])}
        if sim.dim == 3:
            e["omega"] = rel(w, sim.w)
        for k, x in e.items():
            worst[k] = max(worst.get(k, 0.0), x)
        worst["steps_compared"] = worst.get("steps_compared", 0) + 1
    worst["steps"] = worst.get("steps", 0) + 1
    return st, so


def _assert_rounding(worst, bound=1e-9):
    _say("worst relative error vs the FP64 oracle:", worst)
    assert worst.get("steps_compared", 0) >= max(1, int(0.8 * worst["steps"]))
    for k in ("alpha", "beta", "v2", "v1", "omega", "residual"):
        if k in worst:
            assert worst[k] < bound, (k, worst[k])
    assert worst["q1"] <= 1.0


def _tolerance_check(ctx, st, state0, prm, mu, tag, spread_ref=None):
    """A device step judged by the oracle at the step's tolerance (see the module doc)."""
    D = ctx.dim
    gc = ctx.contacts()
    if gc["a"].size == 0:
        return {}
    err, vout = check_with_oracle(D, state0, gc, mu, prm.get("CoR", 0.0))
    v2, w2 = ctx.v2()
    vd = vout.reshape(-1, 6)[:, :3].ravel() if D == 3 else vout
    _say(tag, "device residual", st["error"], "oracle residual of the device impulses", err)
    assert st["succeeded"] and err < prm["tol"] * 1.0001
    assert rel(v2, vd, max(1e-300, np.abs(vd).max())) < 1e-9
    if D == 3:
        assert rel(w2, vout.reshape(-1, 6)[:, 3:]) < 1e-9
    return {"oracle_residual": err}


def _close_at_tolerance(tag, v2_dev, alpha_dev, state0, px, pn, prm):
    """The device's v2 against the oracle's solution at the same tolerance (see the module
    doc): |v2_dev - v2_oracle|_inf <= max(4 sqrt(tol), 10 |v2_natural - v2_colour|_inf)."""
    D = 2 if len(prm["g"]) == 2 else 3
    sims = []
    for order in (1, 0):
        s = oracle.Sim(D, state0["q"], state0["v"], state0["r"], state0["m"], px, pn, w=state0.get("w"),
                       inertia=state0.get("inertia"))
        s.step(order=order, **prm)
        sims.append(s)
    sc, sn = sims
    spread = float(np.abs(sn.v2() - sc.v2()).max())
    d = float(np.abs(v2_dev - sc.v2()).max())
    bound = max(4.0 * np.sqrt(prm["tol"]), 10.0 * spread)
    _say(tag, "|v2 - oracle v2| %.3e m/s (relative %.3e), oracle order spread %.3e, bound %.3e; alpha relative "
         "distance %.3e (oracle order spread %.3e)" % (d, rel(v2_dev, sc.v2()), spread, bound,
                                                          rel(alpha_dev, sc.contacts()["alpha"]),
                                                          rel(sn.contacts()["alpha"], sc.contacts()["alpha"])))
    assert d <= bound


# ----------------------------------------------------------------------------- cfg0
def test_cfg0_lattice_4096_circles_200_steps():
    """The reference's CPU method at its real size: 4096 circles, 200 steps, FP64. The
    device runs all 200 steps; at six of them one step is compared with the oracle from the
    device's own state."""
    s = scenes.ball2d_lattice(m=64, side=10.0, r=0.05, vstd=0.1, seed=1)
    ctx = scx.Context(2, s["r"].size, max_planes=4)
    ctx.set_planes(s["px"], s["pn"])
    ctx.set_bodies(s["q"], s["v"], s["r"], s["m"])
    worst = {}
    sample = {0, 40, 80, 120, 160, 199}
    p = scx.step_params(precision=scx.FP64, **CFG0)
    ok = 0
    for k in range(200):
        if k in sample:
            q, v, _ = ctx.get_state()
            sim = oracle.Sim(2, q, v, s["r"], s["m"], s["px"], s["pn"])
            st, _ = _fp64_step(ctx, sim, CFG0, worst, "cfg0 step %d" % k)
        else:
            st = ctx.step(p)
        ok += int(st["succeeded"])
    ctx.close()
    _say("cfg0: steps at tolerance", ok, "of 200")
    _assert_rounding(worst)


# ----------------------------------------------------------------------------- cfg1
@pytest.fixture(scope="module")
def rsa32k():
    return scenes.ball2d_rsa(n=32768, rmin=0.02, rmax=0.06, phi=0.55, seed=1)


def test_cfg1_rsa_circles_fp64_stepwise(rsa32k):
    """RSA circles (mu 0.3, dt 0.005, symplectic Euler), 32,768 of them: 20 FP64 steps
    in lockstep with the oracle."""
    s = rsa32k
    ctx = scx.Context(2, s["r"].size, max_planes=4)
    ctx.set_planes(s["px"], s["pn"])
    sim = oracle.Sim(2, s["q"], s["v"], s["r"], s["m"], s["px"], s["pn"])
    worst = {}
    for k in range(20):
        _fp64_step(ctx, sim, CFG1, worst, "cfg1-32k step %d" % k)
    ctx.close()
    _assert_rounding(worst)


def test_cfg1_rsa_circles_fp32_polish_meets_tolerance(rsa32k):
    """The benchmark's precision (FP32 sweeps, FP64 polish) on the same packing after 20
    and 40 steps: the oracle's residual of the device impulses meets 1e-6, and the
    distance to the oracle's solution is within ten times its own order spread."""
    s = rsa32k
    ctx = scx.Context(2, s["r"].size, max_planes=4)
    ctx.set_planes(s["px"], s["pn"])
    ctx.set_bodies(s["q"], s["v"], s["r"], s["m"])
    p = scx.step_params(precision=scx.FP32_POLISH, **CFG1)
    for k in range(41):
        if k in (20, 40):
            q, v, _ = ctx.get_state()
            state0 = dict(q=q, v=v, r=s["r"], m=s["m"])
            st = ctx.step(p)
            _tolerance_check(ctx, st, state0, CFG1, CFG1["mu"], "cfg1-32k fp32+polish step %d" % k)
            _close_at_tolerance("cfg1-32k step %d" % k, ctx.v2()[0], ctx.contacts()["alpha"], state0, s["px"],
                                s["pn"], CFG1)
        else:
            ctx.step(p)
    ctx.close()


def _device_vs_oracle_active_set(ctx, q, v, w, s, prm, tag):
    D = ctx.dim
    sim = oracle.Sim(D, q, v, s["r"], s["m"], s["px"], s["pn"], w=w, inertia=s.get("inertia"))
    so = sim.step(order=1, **dict(prm, max_iters=0))
    ctx.set_bodies(q, v, s["r"], s["m"], omega=w, inertia=s.get("inertia"))
    st = ctx.step(scx.step_params(precision=scx.FP64, **dict(prm, max_iters=0)))
    gc, oc = ctx.contacts(), sim.contacts()
    _say(tag, "pairs", st["npairs"], so["npairs"], "contacts", st["ncontacts"], so["ncontacts"], "colours",
         st["ncolours"], so["ncolours"])
    assert st["ncontacts"] == so["ncontacts"] > 0
    assert same_active_set(gc, oc)
    assert np.array_equal(gc["basis"], oc["basis"])
    assert np.array_equal(gc["colour"], oc["colour"])


def test_cfg1_full_size_active_set_bit_exact():
    """1,048,576 circles: the active set (swept boxes, CCD, planes), its order, bases and
    colours against the oracle at step 0 and after 50 device steps at the benchmark's
    precision."""
    s = scenes.ball2d_rsa(n=1048576, rmin=0.02, rmax=0.06, phi=0.55, seed=1)
    ctx = scx.Context(2, s["r"].size, max_planes=4)
    ctx.set_planes(s["px"], s["pn"])
    # step 0 needs a moving state for CCD: take the state after one device step
    ctx.set_bodies(s["q"], s["v"], s["r"], s["m"])
    p = scx.step_params(precision=scx.FP32_POLISH, **CFG1)
    ctx.step(p)
    q, v, _ = ctx.get_state()
    _device_vs_oracle_active_set(ctx, q, v, None, s, CFG1, "cfg1 1M step 1")
    ctx.set_bodies(q, v, s["r"], s["m"])
    for _ in range(49):
        ctx.step(p)
    q, v, _ = ctx.get_state()
    state0 = dict(q=q, v=v, r=s["r"], m=s["m"])
    _device_vs_oracle_active_set(ctx, q, v, None, s, CFG1, "cfg1 1M step 50")
    ctx.set_bodies(q, v, s["r"], s["m"])
    st = ctx.step(p)
    _tolerance_check(ctx, st, state0, CFG1, CFG1["mu"], "cfg1 1M fp32+polish step 50")
    ctx.close()


# ----------------------------------------------------------------------------- cfg2 / cfg3
@pytest.fixture(scope="module")
def pile4k():
    """A poured pile of 4,000 spheres, settled on the device under the transported cache."""
    s = scenes.spheres_pile(n=4000, rmed=0.05, sigma=0.25, seed=3)
    ctx = scx.Context(3, s["r"].size, max_planes=8)
    ctx.set_planes(s["px"], s["pn"])
    ctx.set_bodies(s["q"], s["v"], s["r"], s["m"], inertia=s["inertia"])
    p = scx.step_params(precision=scx.FP32_POLISH, warm_start=scx.WARM_TRANSPORT, **dict(CFG2, max_iters=2000))
    for _ in range(60):
        ctx.step(p)
    q, v, w = ctx.get_state()
    ctx.close()
    return dict(s, q=q, v=v, w=w)


def test_cfg2_pile_fp64_stepwise(pile4k):
    """Plain coloured Gauss-Seidel (cold start, as the reference's 3D geometric map) and
    the transported cache on the settled pile, FP64, in lockstep with the oracle. The
    pile's solves converge slowly (1e-6 needs tens of thousands of sweeps), so every step
    stops at a sweep limit of 2,000 on both sides and the best iterates are compared: over
    more sweeps the rounding differences between the two (FMA contraction, summation
    order) grow, and through each side's own impulse cache they compound from step to
    step."""
    s = pile4k
    ctx = scx.Context(3, s["r"].size, max_planes=8)
    ctx.set_planes(s["px"], s["pn"])
    sim = oracle.Sim(3, s["q"], s["v"], s["r"], s["m"], s["px"], s["pn"], w=s["w"], inertia=s["inertia"])
    worst = {}
    prm = dict(CFG2, max_iters=2000)
    _fp64_step(ctx, sim, prm, worst, "cfg2 cold step")
    # the transported cache: one step on each side fills both caches from the same state
    _fp64_step(ctx, sim, prm, {}, "cfg2 transport fill", warm=scx.WARM_TRANSPORT)
    for k in range(3):
        _fp64_step(ctx, sim, prm, worst, "cfg2 transport step %d" % k, warm=scx.WARM_TRANSPORT)
    ctx.close()
    _assert_rounding(worst)


def test_cfg2_restitution_fp64_stepwise(pile4k):
    """The forcing term of a positive coefficient of restitution (Constraint.cpp:18-34) on
    the pile, FP64, in lockstep with the oracle."""
    s = pile4k
    ctx = scx.Context(3, s["r"].size, max_planes=8)
    ctx.set_planes(s["px"], s["pn"])
    rng = np.random.default_rng(4)
    v = s["v"] + rng.normal(scale=0.05, size=s["v"].size)
    sim = oracle.Sim(3, s["q"], v, s["r"], s["m"], s["px"], s["pn"], w=s["w"], inertia=s["inertia"])
    worst = {}
    for k in range(2):
        _fp64_step(ctx, sim, dict(CFG2, CoR=0.5, max_iters=3000), worst, "cfg2 CoR 0.5 step %d" % k)
    ctx.close()
    _assert_rounding(worst)


def test_cfg2_fp32_polish_and_learned_meet_tolerance(pile4k):
    """The benchmark's precision on the pile with the transported cache and with the
    learned warm start (trained on the device from 8 steps of samples): the oracle's
    residual of the device impulses meets 1e-6; distances to the oracle's solution within
    ten times its own order spread."""
    s = pile4k
    ctx = scx.Context(3, s["r"].size, max_planes=8)
    ctx.set_planes(s["px"], s["pn"])
    ctx.set_bodies(s["q"], s["v"], s["r"], s["m"], omega=s["w"], inertia=s["inertia"])
    tp = scx.step_params(precision=scx.FP32_POLISH, warm_start=scx.WARM_TRANSPORT, **CFG2)
    Xs, Ts = [], []
    for _ in range(8):
        ctx.step(tp)
        X, T = ctx.samples()
        Xs.append(X)
        Ts.append(T)
    from scisim_amd import warmstart
    ctx.set_warm_model(warmstart.init_params(seed=1))
    loss = ctx.train_warm_model(np.concatenate(Xs), np.concatenate(Ts), epochs=6, batch=1024, lr=1e-3, seed=1)
    _say("cfg3 (4k pile) training loss", loss)
    for mode in (scx.WARM_TRANSPORT, scx.WARM_LEARNED):
        q, v, w = ctx.get_state()
        state0 = dict(q=q, v=v, w=w, r=s["r"], m=s["m"], inertia=s["inertia"])
        st = ctx.step(scx.step_params(precision=scx.FP32_POLISH, warm_start=mode, **CFG2))
        _tolerance_check(ctx, st, state0, CFG2, CFG2["mu"], "pile mode %d sweeps %d" % (mode, st["sweeps"]))
        _close_at_tolerance("pile mode %d" % mode, ctx.v2()[0], ctx.contacts()["alpha"], state0, s["px"], s["pn"],
                            dict(CFG2, max_iters=20000))
    ctx.close()


def test_cfg2_full_size_active_set_bit_exact():
    """262,144 spheres poured into a box: active set, order, bases and colours against the
    oracle at the start and after 20 device steps under the transported cache."""
    s = scenes.spheres_pile(n=262144, rmed=0.05, sigma=0.25, seed=1)
    ctx = scx.Context(3, s["r"].size, max_planes=8)
    ctx.set_planes(s["px"], s["pn"])
    ctx.set_bodies(s["q"], s["v"], s["r"], s["m"], inertia=s["inertia"])
    p = scx.step_params(precision=scx.FP32_POLISH, warm_start=scx.WARM_TRANSPORT, **dict(CFG2, max_iters=2000))
    ctx.step(p)
    q, v, w = ctx.get_state()
    _device_vs_oracle_active_set(ctx, q, v, w, s, CFG2, "cfg2 262k step 1")
    ctx.set_bodies(q, v, s["r"], s["m"], omega=w, inertia=s["inertia"])
    for _ in range(19):
        ctx.step(p)
    q, v, w = ctx.get_state()
    _device_vs_oracle_active_set(ctx, q, v, w, s, CFG2, "cfg2 262k step 20")
    ctx.close()


# ----------------------------------------------------------------------------- cfg4
@pytest.fixture(scope="module")
def ensemble8():
    s, body_off, plane_off, mus, counts = scenes.ensemble_piles(nscenes=8, nmin=300, nmax=1500, seed=7)
    return s, body_off, plane_off, mus, counts


def _scene_slice(s, body_off, plane_off, k, q, v, w):
    b0, b1 = int(body_off[k]), int(body_off[k + 1])
    p0, p1 = int(plane_off[k]), int(plane_off[k + 1])
    return dict(q=q[3 * b0:3 * b1], v=v[3 * b0:3 * b1], w=w[3 * b0:3 * b1], r=s["r"][b0:b1], m=s["m"][b0:b1],
                inertia=s["inertia"][b0:b1], px=s["px"][p0:p1], pn=s["pn"][p0:p1]), b0, p0


def _scene_contacts(gc, b0, b1, p0):
    sel = (gc["a"] >= b0) & (gc["a"] < b1)
    out = {k: gc[k][sel] for k in gc}
    out["a"] = out["a"] - b0
    out["b"] = np.where(out["b"] >= 0, out["b"] - b0, -1)
    out["plane"] = np.where(out["plane"] >= 0, out["plane"] - p0, -1)
    return out


def test_cfg4_ensemble_each_scene_vs_oracle(ensemble8):
    """Eight piles of different size and friction stepped together: each scene's active
    set, order and bases bit-exact against the scene stepped alone by the oracle; each
    scene's impulses meet the tolerance by the oracle's residual, and its velocities lie
    within ten times the oracle's order spread of the oracle's solution."""
    s, body_off, plane_off, mus, counts = ensemble8
    ctx = scx.Context(3, s["r"].size, max_planes=64)
    ctx.set_planes(s["px"], s["pn"])
    ctx.set_bodies(s["q"], s["v"], s["r"], s["m"], inertia=s["inertia"])
    ctx.set_scenes(body_off, plane_off, mus)
    tp = scx.step_params(precision=scx.FP32_POLISH, warm_start=scx.WARM_TRANSPORT, **dict(CFG2, max_iters=2000))
    for _ in range(30):
        ctx.step(tp)
    q, v, w = ctx.get_state()
    st = ctx.step(scx.step_params(precision=scx.FP32_POLISH, **CFG2))
    gc = ctx.contacts()
    v2, w2 = ctx.v2()
    cnt, err, ssw = ctx.scene_stats()
    ctx.close()
    for k in range(8):
        sc, b0, p0 = _scene_slice(s, body_off, plane_off, k, q, v, w)
        b1 = int(body_off[k + 1])
        mine = _scene_contacts(gc, b0, b1, p0)
        alone = oracle.Sim(3, sc["q"], sc["v"], sc["r"], sc["m"], sc["px"], sc["pn"], w=sc["w"], inertia=sc["inertia"])
        so = alone.step(order=1, **dict(CFG2, mu=float(mus[k]), max_iters=0))
        oc = alone.contacts()
        assert cnt[k] == so["ncontacts"] == mine["a"].size
        assert same_active_set(mine, oc)
        assert np.array_equal(mine["basis"], oc["basis"])
        e, vout = check_with_oracle(3, sc, mine, float(mus[k]))
        _say("scene", k, "bodies", b1 - b0, "mu %.3f" % mus[k], "contacts", cnt[k], "sweeps", ssw[k],
             "device residual", err[k], "oracle residual of the device impulses", e)
        assert e < CFG2["tol"] * 1.0001
        assert rel(v2[3 * b0:3 * b1], vout.reshape(-1, 6)[:, :3]) < 1e-9
        _close_at_tolerance("scene %d" % k, v2[3 * b0:3 * b1], mine["alpha"], sc, sc["px"], sc["pn"],
                            dict(CFG2, mu=float(mus[k])))


@pytest.mark.parametrize("mode", [scx.WARM_NORMAL, scx.WARM_TRANSPORT])
def test_cfg4_ensemble_warm_start_finds_every_cached_contact(ensemble8, mode):
    """Each scene of an ensemble receives the warm start of every contact that the previous
    step cached, as the scene alone would: the cached normal impulse returns (NORMAL
    restores alpha along the new normal, the transported cache rotates the world impulse
    with the normal), plane contacts of every scene included."""
    s, body_off, plane_off, mus, counts = ensemble8
    ctx = scx.Context(3, s["r"].size, max_planes=64)
    ctx.set_planes(s["px"], s["pn"])
    ctx.set_bodies(s["q"], s["v"], s["r"], s["m"], inertia=s["inertia"])
    ctx.set_scenes(body_off, plane_off, mus)
    for _ in range(10):
        ctx.step(scx.step_params(precision=scx.FP32_POLISH, warm_start=mode, **dict(CFG2, max_iters=2000)))
    old = ctx.contacts()
    # no sweep: the step's impulses are its warm start
    ctx.step(scx.step_params(precision=scx.FP64, warm_start=mode, **dict(CFG2, max_iters=0)))
    new = ctx.contacts()
    ctx.close()
    key = lambda c, k: (int(c["a"][k]), int(c["b"][k]), int(c["plane"][k]))
    prev = {key(old, k): k for k in range(old["a"].size)}
    hits = planes = 0
    for k in range(new["a"].size):
        j = prev.get(key(new, k))
        if j is None:
            continue
        hits += 1
        if new["b"][k] < 0:
            planes += 1
        # NORMAL restores alpha along the new normal; the transport rotation keeps the
        # normal component of the world impulse
        assert abs(new["alpha"][k] - old["alpha"][j]) <= 1e-12 * abs(old["alpha"][j]), (k, new["alpha"][k],
                                                                                        old["alpha"][j])
    _say("warm start mode", mode, "cached contacts found", hits, "of them plane contacts", planes)
    assert planes > 100 and hits > 1000


# ----------------------------------------------------------------------------- errors
def test_colour_overflow_is_reported_before_the_state_changes():
    """A hub circle touching 70 others needs more colours than the sweeps support:
    SCX_E_COLOURS from scx_step and scx_run, with the state left as it was."""
    R, r, k = 0.25, 0.01, 70
    ang = 2 * np.pi * np.arange(k) / k
    q = np.r_[[0.0, 0.0], np.c_[(R + r) * 0.9999 * np.cos(ang), (R + r) * 0.9999 * np.sin(ang)].ravel()]
    rad = np.r_[R, np.full(k, r)]
    m = np.pi * rad * rad
    ctx = scx.Context(2, k + 1, max_planes=4)
    ctx.set_planes(np.zeros((0, 2)), np.zeros((0, 2)))
    ctx.set_bodies(q, np.zeros(2 * (k + 1)), rad, m)
    prm = scx.step_params(precision=scx.FP64, **dict(CFG0, g=[0.0, 0.0]))
    for call in (lambda: ctx.step(prm), lambda: ctx.run(prm, 3)):
        with pytest.raises(scx.ScxError) as ei:
            call()
        assert "(-5)" in str(ei.value)
        q1, v1, _ = ctx.get_state()
        assert np.array_equal(q1, q) and np.array_equal(v1, np.zeros_like(v1))
    ctx.close()


def test_friction_solve_rejects_invalid_bodies():
    ctx = scx.Context(3, 16, max_planes=4)
    n = np.array([0.0, 1.0, 0.0])
    basis = np.r_[n, 1.0, 0.0, 0.0, 0.0, 0.0, -1.0]
    kw = dict(arm0=[0.0, -0.1, 0.0], arm1=[0.0, 0.0, 0.0])
    with pytest.raises(scx.ScxError):  # 3D without inertia
        ctx.friction_solve([1.0], [0, -1, 0, 0, 0, 0][:3], [0], [-1], basis, [0.5], omega0=np.zeros(3), **kw)
    with pytest.raises(scx.ScxError):  # zero mass
        ctx.friction_solve([0.0], [0, -1, 0], [0], [-1], basis, [0.5], inertia=[1.0], omega0=np.zeros(3), **kw)
    with pytest.raises(scx.ScxError):  # zero inertia
        ctx.friction_solve([1.0], [0, -1, 0], [0], [-1], basis, [0.5], inertia=[0.0], omega0=np.zeros(3), **kw)
    res = ctx.friction_solve([1.0], [0, -1, 0], [0], [-1], basis, [0.5], inertia=[0.004], omega0=np.zeros(3), **kw)
    assert res["succeeded"] and abs(res["alpha"][0] - 1.0) < 1e-15
    ctx.close()

"""GPU parity: the HIP library through the C ABI against the FP64 CPU oracle.

Bit-exact: pair lists (reference fixtures and generated scenes), active sets and
their order, contact bases, colours, unconstrained flows. Floating point: impulses
and post-step velocities at the same residual tolerance, compared with the oracle in
the same colour order (rounding-level agreement) and in the reference's natural
order (agreement bounded by the residual tolerance); the tolerances are stated in
each test.
"""

import os
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


@pytest.fixture(scope="module")
def ctx2():
    c = scx.Context(2, 40000, max_planes=8)
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx3():
    c = scx.Context(3, 40000, max_planes=8)
    yield c
    c.close()


# ----------------------------------------------------------------------------- broad phase
@pytest.mark.parametrize("name", ["test_00_data", "test_01_data", "test_02_data"])
def test_overlaps_reference_fixtures(ctx2, golden_dir, name):
    a = np.load(os.path.join(golden_dir, "aabb_fixtures.npz"))[name]
    _say("fixture", name, a.shape)
    got = ctx2.potential_overlaps(a[:, :2], a[:, 2:])
    ref = oracle.all_pairs(2, a[:, :2], a[:, 2:])
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("seed", [0, 1])
def test_overlaps_random_3d(ctx3, seed):
    rng = np.random.default_rng(seed)
    n = 20000
    c = rng.uniform(-3, 3, (n, 3))
    h = np.exp(rng.normal(np.log(0.05), 0.5, (n, 3)))
    mn, mx = c - h, c + h
    _say("random 3d", seed)
    assert np.array_equal(ctx3.potential_overlaps(mn, mx), oracle.grid_pairs(3, mn, mx))


def test_overlaps_edge_cases(ctx2):
    _say("edge cases")
    # single box: no pairs
    assert ctx2.potential_overlaps(np.array([[0.0, 0.0]]), np.array([[1.0, 1.0]])).shape == (0, 2)
    # identical boxes: every pair
    mn = np.zeros((60, 2))
    mx = np.ones((60, 2))
    got = ctx2.potential_overlaps(mn, mx)
    assert got.shape[0] == 60 * 59 // 2 and np.array_equal(got, oracle.all_pairs(2, mn, mx))
    # touching faces overlap; a gap of one ulp does not
    mn = np.array([[0.0, 0.0], [1.0, 0.0], [np.nextafter(2.0, 3.0), 0.0]])
    mx = np.array([[1.0, 1.0], [2.0, 1.0], [3.0, 1.0]])
    assert ctx2.potential_overlaps(mn, mx).tolist() == [[0, 1]]
    # one huge box among small ones
    rng = np.random.default_rng(5)
    c = rng.uniform(-10, 10, (3000, 2))
    mn = c - 0.05
    mx = c + 0.05
    mn[17] = [-20, -20]
    mx[17] = [20, 20]
    assert np.array_equal(ctx2.potential_overlaps(mn, mx), oracle.all_pairs(2, mn, mx))
    # ragged sizes over six orders of magnitude
    h = 10.0 ** rng.uniform(-4, 1, (3000, 2))
    assert np.array_equal(ctx2.potential_overlaps(c - h, c + h), oracle.all_pairs(2, c - h, c + h))


# ----------------------------------------------------------------------------- flows
@pytest.mark.parametrize("dim,integrator", [(2, 0), (2, 1), (3, 1)])
def test_unconstrained_flow_bitexact(ctx2, ctx3, dim, integrator):
    rng = np.random.default_rng(dim * 10 + integrator)
    n = 5000
    m = rng.uniform(0.1, 3.0, n)
    q0 = rng.normal(size=dim * n)
    v0 = rng.normal(size=dim * n)
    g = [0.1, -9.81, 0.3][:dim]
    ctx = ctx2 if dim == 2 else ctx3
    q1, v1 = ctx.unconstrained_flow(integrator, m, g, 0.005, q0, v0)
    oq, ov = oracle.umap_flow(dim, integrator, m, g + [0.0] * (3 - dim), 0.005, q0, v0)
    assert np.array_equal(q1, oq) and np.array_equal(v1, ov)


# ----------------------------------------------------------------------------- full steps
def _params(dim, **kw):
    d = dict(dt=0.01, g=[0.0, -9.81] if dim == 2 else [0.0, -9.81, 0.0], mu=0.5, CoR=0.0, tol=1e-12,
             max_iters=5000, eval_every=1, integrator=0 if dim == 2 else 1)
    d.update(kw)
    return d


def _compare_step(ctx, sim, prm, precision=scx.FP64, tag=""):
    """One step of both from the same state; returns error metrics."""
    D = sim.dim
    ctx.set_bodies(sim.q, sim.v, sim.r, sim.m, omega=sim.w, inertia=sim.inertia)
    st = ctx.step(scx.step_params(precision=precision, **prm))
    so = sim.step(order=1, **prm)
    gc = ctx.contacts()
    oc = sim.contacts()
    _say(tag, "contacts", st["ncontacts"], "colours", st["ncolours"], "sweeps", st["sweeps"], so["sweeps"])
    # active set: identical and identically ordered
    assert st["ncontacts"] == so["ncontacts"]
    assert np.array_equal(gc["a"], oc["a"]) and np.array_equal(gc["b"], oc["b"])
    assert np.array_equal(gc["plane"], oc["plane"])
    assert np.array_equal(gc["basis"], oc["basis"])
    assert np.array_equal(gc["colour"], oc["colour"])
    q, v, w = ctx.get_state()
    scale_v = max(1e-30, np.abs(sim.v).max())
    amax = max(1e-30, np.abs(oc["alpha"]).max()) if oc["alpha"].size else 1.0
    return {
        "alpha": np.abs(gc["alpha"] - oc["alpha"]).max() / amax if oc["alpha"].size else 0.0,
        "beta": np.abs(gc["beta"] - oc["beta"]).max() / amax if oc["beta"].size else 0.0,
        "v": np.abs(v - sim.v).max() / scale_v,
        "q": np.abs(q - sim.q).max(),
        "sweeps": (st["sweeps"], so["sweeps"]),
        "error": (st["error"], so["error"]),
        "ncontacts": st["ncontacts"],
    }


def test_step_2d_lattice_matches_oracle(ctx2):
    s = scenes.ball2d_lattice(m=48, side=7.5, r=0.05, vstd=0.1, seed=1)
    sim = oracle.Sim(2, s["q"], s["v"], s["r"], s["m"], s["px"], s["pn"])
    ctx2.set_planes(s["px"], s["pn"])
    prm = _params(2, dt=0.01, mu=0.5)
    # warm up on the oracle so that a pile forms, then compare step by step
    sim.run(60, **prm)
    worst = {}
    for k in range(8):
        e = _compare_step(ctx2, sim, prm, tag="2d step %d" % k)
        for key in ("alpha", "beta", "v", "q"):
            worst[key] = max(worst.get(key, 0.0), e[key])
        assert abs(e["sweeps"][0] - e["sweeps"][1]) <= 1
    _say("2d worst", worst)
    # same colour order, FP64: rounding-level agreement
    assert worst["alpha"] < 1e-9 and worst["beta"] < 1e-9 and worst["v"] < 1e-9 and worst["q"] < 1e-12


def _spheres_in_box(n, seed):
    s = scenes.spheres_rsa(n=n, phi=0.35, seed=seed)
    return s


def test_step_3d_spheres_matches_oracle(ctx3):
    s = _spheres_in_box(3000, 2)
    sim = oracle.Sim(3, s["q"], s["v"], s["r"], s["m"], s["px"], s["pn"], inertia=s["inertia"])
    ctx3.set_planes(s["px"], s["pn"])
    prm = _params(3, dt=0.005, mu=0.4)
    sim.run(80, **prm)
    worst = {}
    for k in range(6):
        e = _compare_step(ctx3, sim, prm, tag="3d step %d" % k)
        for key in ("alpha", "beta", "v", "q"):
            worst[key] = max(worst.get(key, 0.0), e[key])
        assert abs(e["sweeps"][0] - e["sweeps"][1]) <= 1
    _say("3d worst", worst)
    assert worst["alpha"] < 1e-8 and worst["beta"] < 1e-8 and worst["v"] < 1e-8 and worst["q"] < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_fp32_sweeps_with_fp64_polish(ctx2, ctx3, dim):
    if dim == 2:
        s = scenes.ball2d_lattice(m=48, side=7.5, r=0.05, vstd=0.1, seed=3)
        sim = oracle.Sim(2, s["q"], s["v"], s["r"], s["m"], s["px"], s["pn"])
        # a loose pile, where Gauss-Seidel reaches a tight tolerance (jammed piles stall,
        # in every contact order, as projected GS does on frictional networks)
        ctx, prm = ctx2, _params(2, dt=0.01, mu=0.5, tol=1e-10, max_iters=20000)
        sim.run(25, **prm)
    else:
        s = _spheres_in_box(3000, 4)
        sim = oracle.Sim(3, s["q"], s["v"], s["r"], s["m"], s["px"], s["pn"], inertia=s["inertia"])
        ctx, prm = ctx3, _params(3, dt=0.005, mu=0.4, tol=1e-10, max_iters=20000)
        sim.run(40, **prm)
    ctx.set_planes(s["px"], s["pn"])
    ctx.set_bodies(sim.q, sim.v, sim.r, sim.m, omega=sim.w, inertia=sim.inertia)
    st = ctx.step(scx.step_params(precision=scx.FP32_POLISH, **prm))
    _say("fp32+polish", dim, st)
    so = sim.step(order=1, **prm)
    assert st["succeeded"] and st["error"] < prm["tol"]
    gc, oc = ctx.contacts(), sim.contacts()
    assert np.array_equal(gc["a"], oc["a"]) and np.array_equal(gc["b"], oc["b"])
    # both solutions satisfy the tolerance; their distance is bounded by it
    q, v, w = ctx.get_state()
    dv = np.abs(v - sim.v).max() / max(1e-30, np.abs(sim.v).max())
    _say("fp32+polish rel v err", dv)
    assert dv < 1e-3


def test_natural_order_oracle_agreement(ctx2):
    """GPU colour order vs the reference's natural order at the same tolerance.

    Both orders converge to a Coulomb solution within the residual tolerance; on a
    loose pile (unique solution) the post-step velocities then agree to a bound set by
    the tolerance: ||FB|| < 1e-8 here, velocities O(1), so a relative 1e-5 bound."""
    s = scenes.ball2d_lattice(m=48, side=7.5, r=0.05, vstd=0.1, seed=5)
    sim = oracle.Sim(2, s["q"], s["v"], s["r"], s["m"], s["px"], s["pn"])
    prm = _params(2, dt=0.01, mu=0.5, tol=1e-16, max_iters=20000)
    sim.run(25, **prm)
    ctx2.set_planes(s["px"], s["pn"])
    ctx2.set_bodies(sim.q, sim.v, sim.r, sim.m)
    st = ctx2.step(scx.step_params(**prm))
    so = sim.step(order=0, **prm)
    q, v, w = ctx2.get_state()
    dv = np.abs(v - sim.v).max() / np.abs(sim.v).max()
    _say("natural-order rel v err", dv, st, so)
    assert st["succeeded"] and so["succeeded"]
    assert dv < 1e-5


def test_empty_active_set_keeps_predictor(ctx2):
    n = 10
    q = np.c_[np.arange(n) * 1.0, np.zeros(n)].ravel()
    v = np.zeros(2 * n)
    ctx2.set_planes(np.zeros((0, 2)), np.zeros((0, 2)))
    ctx2.set_bodies(q, v, np.full(n, 0.1), np.ones(n))
    st = ctx2.step(scx.step_params(**_params(2)))
    assert st["ncontacts"] == 0
    q1, v1, _ = ctx2.get_state()
    oq, ov = oracle.umap_flow(2, 0, np.ones(n), [0.0, -9.81, 0.0], 0.01, q, v)
    assert np.array_equal(q1, oq) and np.array_equal(v1, ov)


@pytest.mark.parametrize("dim,mode", [(2, scx.WARM_CACHE), (2, scx.WARM_NORMAL), (3, scx.WARM_NORMAL),
                                      (3, scx.WARM_TRANSPORT)])
def test_impulse_cache_warm_start_matches_oracle(dim, mode):
    """The impulse cache (ImpulsesToCache) carried between steps on the device gives the
    oracle's warm starts. At a tolerance neither side reaches (fixed sweep count) the
    impulses agree to rounding; at 1e-6 the warm-started solves stop within two sweeps of
    the oracle's and take fewer sweeps than cold starts from the same states."""
    ctx = scx.Context(dim, 40000, max_planes=8)  # its own, empty cache
    if dim == 2:
        s = scenes.ball2d_lattice(m=48, side=7.5, r=0.05, vstd=0.1, seed=5)
        sim = oracle.Sim(2, s["q"], s["v"], s["r"], s["m"], s["px"], s["pn"])
        prm = _params(2, dt=0.01, mu=0.5, tol=1e-10, max_iters=2000, warm_start=mode)
        warmup = 60
    else:
        s = _spheres_in_box(3000, 4)
        sim = oracle.Sim(3, s["q"], s["v"], s["r"], s["m"], s["px"], s["pn"], inertia=s["inertia"])
        prm = _params(3, dt=0.005, mu=0.4, tol=1e-10, max_iters=2000, warm_start=mode)
        warmup = 80
    ctx.set_planes(s["px"], s["pn"])
    sim.run(warmup, **dict(prm, warm_start=0))
    # one step on each side fills both caches from the same state
    _compare_step(ctx, sim, prm, tag="warm fill")
    worst = {}
    for k in range(3):
        e = _compare_step(ctx, sim, prm, tag="warm %dd step %d" % (dim, k))
        for key in ("alpha", "beta", "v", "q"):
            worst[key] = max(worst.get(key, 0.0), e[key])
        assert abs(e["sweeps"][0] - e["sweeps"][1]) <= 1
    _say("warm", dim, mode, worst)
    assert worst["alpha"] < 1e-8 and worst["beta"] < 1e-8 and worst["v"] < 1e-8 and worst["q"] < 1e-11
    conv = dict(prm, tol=1e-6, max_iters=5000)
    cold = dict(conv, warm_start=0)
    warm_sweeps, oracle_sweeps, cold_sweeps = [], [], []
    for k in range(4):
        probe = oracle.Sim(dim, sim.q, sim.v, sim.r, sim.m, sim.px, sim.pn, w=sim.w, inertia=sim.inertia)
        cold_sweeps.append(probe.step(order=1, **cold)["sweeps"])
        e = _compare_step(ctx, sim, conv, tag="warm %dd tol 1e-6 step %d" % (dim, k))
        warm_sweeps.append(e["sweeps"][0])
        oracle_sweeps.append(e["sweeps"][1])
    _say("sweeps warm", warm_sweeps, "oracle warm", oracle_sweeps, "cold", cold_sweeps)
    ctx.close()
    assert all(abs(a - b) <= 2 for a, b in zip(warm_sweeps, oracle_sweeps))
    assert sum(warm_sweeps) < sum(cold_sweeps)


def test_normal_and_friction_cache_is_rejected_in_3d(ctx3):
    s = _spheres_in_box(200, 1)
    ctx3.set_planes(s["px"], s["pn"])
    ctx3.set_bodies(s["q"], s["v"], s["r"], s["m"], inertia=s["inertia"])
    with pytest.raises(scx.ScxError):
        ctx3.step(scx.step_params(**_params(3, warm_start=scx.WARM_CACHE)))


@pytest.mark.parametrize("dim", [2, 3])
def test_multi_workgroup_barrier_paths_match_oracle(monkeypatch, dim):
    """The persistent colouring and sweep kernels on 16 workgroups (device-wide barriers
    between colours and rounds) give the oracle's colours and rounding-level impulses, as
    the single-workgroup launches that small problems get by default."""
    monkeypatch.setenv("SCX_FORCE_PERSISTENT_BLOCKS", "16")
    ctx = scx.Context(dim, 40000, max_planes=8)
    if dim == 2:
        s = scenes.ball2d_lattice(m=48, side=7.5, r=0.05, vstd=0.1, seed=11)
        sim = oracle.Sim(2, s["q"], s["v"], s["r"], s["m"], s["px"], s["pn"])
        prm = _params(2, dt=0.01, mu=0.5, max_iters=1000)
        warmup = 60
    else:
        s = _spheres_in_box(3000, 6)
        sim = oracle.Sim(3, s["q"], s["v"], s["r"], s["m"], s["px"], s["pn"], inertia=s["inertia"])
        prm = _params(3, dt=0.005, mu=0.4, max_iters=1000)
        warmup = 80
    ctx.set_planes(s["px"], s["pn"])
    sim.run(warmup, **prm)
    worst = {}
    for k in range(3):
        e = _compare_step(ctx, sim, prm, tag="16 workgroups %dd step %d" % (dim, k))
        for key in ("alpha", "beta", "v", "q"):
            worst[key] = max(worst.get(key, 0.0), e[key])
        assert abs(e["sweeps"][0] - e["sweeps"][1]) <= 1
    # FP32 sweeps + FP64 polish on the same launch shape
    ctx.set_bodies(sim.q, sim.v, sim.r, sim.m, omega=sim.w, inertia=sim.inertia)
    st = ctx.step(scx.step_params(precision=scx.FP32_POLISH, **dict(prm, tol=1e-6, max_iters=5000)))
    ctx.close()
    _say("16 workgroups", dim, worst, "fp32+polish", st)
    assert worst["alpha"] < 1e-8 and worst["beta"] < 1e-8 and worst["v"] < 1e-8 and worst["q"] < 1e-11
    assert st["ncontacts"] > 0 and np.isfinite(st["error"])


@pytest.mark.parametrize("dim", [2, 3])
def test_a_step_without_caching_leaves_the_cache_on_the_device(dim):
    """As the oracle: a cold step (NONE) leaves the device cache as it was, so the next
    caching step warm-starts a resting body with no sweep; an empty active set empties it."""
    ctx = scx.Context(dim, 16, max_planes=4)
    g = [0.0, -9.81, 0.0][:dim]
    q = np.zeros(dim)
    q[1] = 0.1
    ctx.set_planes(np.zeros(dim), np.eye(dim)[1])
    ctx.set_bodies(q, np.zeros(dim), [0.1], [2.0])

    def step(mode):
        return ctx.step(scx.step_params(**_params(dim, mu=0.5, tol=1e-12, warm_start=mode)))

    for _ in range(3):
        step(scx.WARM_NORMAL)
    assert step(scx.WARM_NONE)["sweeps"] >= 1
    assert step(scx.WARM_NORMAL)["sweeps"] == 0
    lifted = q.copy()
    lifted[1] = 5.0
    ctx.set_bodies(lifted, np.zeros(dim), [0.1], [2.0])  # same body count: the cache is kept
    assert step(scx.WARM_NONE)["ncontacts"] == 0         # empty active set: the cache is emptied
    v = np.zeros(dim)
    v[1] = -9.81 * 0.01
    ctx.set_bodies(q, v, [0.1], [2.0])
    assert step(scx.WARM_NORMAL)["sweeps"] >= 1
    assert step(scx.WARM_NORMAL)["sweeps"] == 0
    ctx.close()

"""Pins the oracle's symplectic-Euler and stabilized staggerings
(scisim/ConstrainedMaps/SymplecticEulerImpactFrictionMap.cpp:201-407,
StabilizedImpactFrictionMap.cpp:48-183) by what the mathematics fixes:

* symplectic Euler puts the force increment into the solve (nrel = N^T vdelta,
  drel = D^T vdelta), so a body resting on an incline follows the continuous solution
  from the first step on: at rest when static friction holds (2D, mu >= tan theta),
  rolling a = 5/7 g sin(theta) (3D, mu >= 2/7 tan theta), else sliding
  a = g (sin(theta) - mu cos(theta)); a dropped drel would let it slide freely;
* a stack resting on a floor keeps v = 0 exactly and carries k m g dt in its k-th
  contact from the top from the first step (nrel of body pairs and planes);
* stabilization adds min(0, g(q0)) / dt to the normal forcing: a body sunk by delta
  into the floor leaves with normal speed delta / dt and ends the step at zero gap,
  unless delta is below the penetration threshold;
* contact sets among free bodies change the total momentum by exactly the gravity
  impulse sum(m) g dt (the reference's momentum check, :336-347);
* the stabilized map solves on the predictor velocity: a body at rest on a floor stays
  at rest, and a free body follows the unconstrained map (:66-78);
* empty active sets give the symplectic Euler closed form q0 + dt (v0 + g dt).
"""

import numpy as np
import pytest

import oracle

G = 9.81


def _incline3(theta, azimuth):
    n = np.array([-np.sin(theta), np.cos(theta), 0.0])
    ca, sa = np.cos(azimuth), np.sin(azimuth)
    Ry = np.array([[ca, 0, sa], [0, 1, 0], [-sa, 0, ca]])
    n = Ry @ n
    g = np.array([0.0, -G, 0.0])
    gt = g - (g @ n) * n
    return n, gt / np.linalg.norm(gt)


@pytest.mark.parametrize("staggering", [1, 2])
@pytest.mark.parametrize("theta_deg,mu,azimuth", [(20, 0.5, 0.0), (35, 0.3, 0.7), (20, 0.05, 1.9),
                                                  (35, 0.1, -0.4), (10, 0.0, 2.5)])
def test_sphere_on_incline_follows_the_closed_form_from_the_first_step(staggering, theta_deg, mu, azimuth):
    theta = np.radians(theta_deg)
    n, d = _incline3(theta, azimuth)
    rad, m, dt = 0.1, 2.0, 0.005
    sim = oracle.Sim(3, rad * n, np.zeros(3), [rad], [m], np.zeros(3), n)
    rolling = mu >= 2.0 / 7.0 * np.tan(theta)
    a = 5.0 / 7.0 * G * np.sin(theta) if rolling else G * (np.sin(theta) - mu * np.cos(theta))
    spin = a / rad if rolling else 5.0 * mu * G * np.cos(theta) / (2.0 * rad)
    axis = np.cross(n, d)
    for k in range(1, 21):
        st = sim.step(dt=dt, g=[0.0, -G, 0.0], mu=mu, tol=1e-26, max_iters=100, integrator=1,
                      staggering=staggering)
        assert st["ncontacts"] == 1
        # velocity and spin after k steps: k a dt along the slope, k spin dt about n x d
        assert abs(sim.v @ d - k * a * dt) < 1e-12 * max(1.0, a)
        assert np.linalg.norm(sim.v - (sim.v @ d) * d) < 1e-12
        assert abs(sim.w @ axis - k * spin * dt) < 1e-11 * max(1.0, spin)
        assert np.linalg.norm(sim.w - (sim.w @ axis) * axis) < 1e-11
    # the sphere stays on the plane
    assert abs(sim.q @ n - rad) < 1e-12


@pytest.mark.parametrize("staggering", [1, 2])
@pytest.mark.parametrize("theta_deg,mu", [(20, 0.5), (30, 0.6), (30, 0.2), (45, 0.5), (15, 0.0)])
def test_circle_on_incline_sticks_or_slides_from_the_first_step(staggering, theta_deg, mu):
    theta = np.radians(theta_deg)
    n = np.array([-np.sin(theta), np.cos(theta)])
    d = np.array([-np.cos(theta), -np.sin(theta)])  # downhill
    rad, m, dt = 0.05, 1.3, 0.01
    sim = oracle.Sim(2, rad * n, np.zeros(2), [rad], [m], np.zeros(2), n)
    a = 0.0 if mu >= np.tan(theta) else G * (np.sin(theta) - mu * np.cos(theta))
    for k in range(1, 31):
        st = sim.step(dt=dt, g=[0.0, -G], mu=mu, tol=1e-26, max_iters=100, integrator=0, staggering=staggering)
        assert st["ncontacts"] == 1
        assert abs(sim.v @ d - k * a * dt) < 1e-12
        assert abs(sim.v @ n) < 1e-12
        c = sim.contacts()
        # the normal impulse carries the normal weight component m g cos(theta) dt
        assert abs(c["alpha"][0] - m * G * np.cos(theta) * dt) < 1e-12
    if a == 0.0:
        assert np.allclose(sim.q, rad * n, rtol=0, atol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_symplectic_euler_stack_holds_its_weight_from_the_first_step(dim):
    nb, rad, m, dt = 6, 0.05, 0.7, 0.005
    q = np.zeros((nb, dim))
    # slightly overlapping, so that every pair is active (2D: CCD of bodies that do not move
    # relative to each other needs an overlap at q0)
    q[:, 1] = rad * (1 - 1e-6) + 2 * rad * (1 - 1e-6) * np.arange(nb)
    up = np.zeros(dim)
    up[1] = 1.0
    g = [0.0, -G, 0.0][:dim]
    sim = oracle.Sim(dim, q.ravel(), np.zeros(dim * nb), np.full(nb, rad), np.full(nb, m), np.zeros(dim), up)
    for _ in range(3):
        st = sim.step(dt=dt, g=g, mu=0.3, tol=1e-28, max_iters=100000, integrator=0 if dim == 2 else 1,
                      staggering=1, order=0)
        assert st["succeeded"] and st["ncontacts"] == nb
        assert np.abs(sim.v).max() < 1e-13
        c = sim.contacts()
        w = m * G * dt
        expect = np.r_[(nb - 1 - np.arange(nb - 1)) * w, nb * w]
        assert np.allclose(c["alpha"], expect, rtol=1e-10, atol=0)
        assert np.abs(c["beta"]).max() < 1e-12 * w
    assert np.allclose(sim.q, q.ravel(), rtol=0, atol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("delta,threshold", [(2e-3, 1e-4), (2e-3, 0.0), (5e-5, 1e-4)])
def test_stabilization_pushes_out_sunk_bodies(dim, delta, threshold):
    rad, m, dt = 0.1, 1.0, 0.01
    up = np.zeros(dim)
    up[1] = 1.0
    q0 = (rad - delta) * up
    g = [0.0, -G, 0.0][:dim]
    sim = oracle.Sim(dim, q0, np.zeros(dim), [rad], [m], np.zeros(dim), up)
    st = sim.step(dt=dt, g=g, mu=0.5, tol=1e-28, max_iters=100, integrator=0, staggering=1, stabilize=1,
                  penetration_threshold=threshold)
    assert st["ncontacts"] == 1
    pushed = delta >= threshold
    expect_v = delta / dt if pushed else 0.0
    assert abs(sim.v[1] - expect_v) < 1e-12
    assert np.abs(np.delete(sim.v, 1)).max() < 1e-15
    # a pushed body ends the step at zero gap; an unpushed one where it was
    assert abs(sim.q[1] - (rad if pushed else rad - delta)) < 1e-13
    # without stabilization the sunk body is held, not pushed
    sim2 = oracle.Sim(dim, q0, np.zeros(dim), [rad], [m], np.zeros(dim), up)
    sim2.step(dt=dt, g=g, mu=0.5, tol=1e-28, max_iters=100, integrator=0, staggering=1, stabilize=0)
    assert abs(sim2.v[1]) < 1e-12


def _cluster(dim, seed, nb=12):
    rng = np.random.default_rng(seed)
    # a loose cloud of overlapping or nearly touching bodies, no planes
    x = rng.uniform(-0.12, 0.12, (nb, dim))
    v = rng.normal(0.0, 0.5, (nb, dim))
    r = rng.uniform(0.04, 0.06, nb)
    m = rng.uniform(0.5, 2.0, nb)
    return x, v, r, m


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_symplectic_euler_momentum_changes_by_the_gravity_impulse(dim, seed):
    x, v, r, m = _cluster(dim, seed)
    g = np.array([0.0, -G, 0.0][:dim])
    dt = 0.01
    w = np.random.default_rng(seed + 10).normal(0, 1.0, 3 * len(m)) if dim == 3 else None
    sim = oracle.Sim(dim, x.ravel(), v.ravel(), r, m, np.zeros((0, dim)), np.zeros((0, dim)), w=w)
    p0 = (m[:, None] * v).sum(0)
    st = sim.step(dt=dt, g=list(g), mu=0.4, tol=1e-24, max_iters=200000, integrator=0 if dim == 2 else 1,
                  staggering=1)
    assert st["ncontacts"] > 3 and st["succeeded"]
    p1 = (m[:, None] * sim.v.reshape(-1, dim)).sum(0)
    assert np.allclose(p1, p0 + m.sum() * g * dt, rtol=0, atol=1e-12)
    if dim == 3:
        # angular momentum about the origin: sum x0 x (m v) + I omega, conserved by the
        # impulses (gravity adds sum x0 x m g dt); the reference checks it on v1 before
        # the force increment is added (:336-347)
        I = 0.4 * m * r * r
        L0 = np.cross(x, m[:, None] * v).sum(0) + (I[:, None] * w.reshape(-1, 3)).sum(0)
        v1 = sim.v.reshape(-1, 3) - g * dt
        L1 = np.cross(x, m[:, None] * v1).sum(0) + (I[:, None] * sim.w.reshape(-1, 3)).sum(0)
        assert np.allclose(L1, L0, rtol=0, atol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_stabilized_map_keeps_a_resting_body_at_rest(dim):
    rad, m, dt = 0.1, 1.0, 0.01
    up = np.zeros(dim)
    up[1] = 1.0
    g = np.array([0.0, -G, 0.0][:dim])
    integ = 1  # Verlet: its predictor differs from q0 + dt v1, so the two updates are told apart
    q0 = np.r_[rad * up, (5.0 * up + np.r_[1.0, np.zeros(dim - 1)])]
    v0 = np.r_[np.zeros(dim), np.r_[0.3, 0.2, -0.1][:dim]]
    sim = oracle.Sim(dim, q0, v0, [rad, rad], [m, m], np.zeros(dim), up)
    for _ in range(5):
        st = sim.step(dt=dt, g=list(g), mu=0.5, tol=1e-28, max_iters=100, integrator=integ, staggering=2)
        assert st["ncontacts"] == 1
    assert np.abs(sim.v[:dim]).max() < 1e-13
    assert np.allclose(sim.q[:dim], rad * up, rtol=0, atol=1e-14)
    # with a non-empty active set every body ends on q0 + dt v1, v1 the predictor velocity
    # v0 + g dt (StabilizedImpactFrictionMap.cpp:181): after k steps v0 + k g dt and
    # q0 + k dt v0 + k (k + 1) / 2 g dt^2
    k = 5
    assert np.allclose(sim.v[dim:], v0[dim:] + k * g * dt, rtol=0, atol=1e-13)
    assert np.allclose(sim.q[dim:], q0[dim:] + k * dt * v0[dim:] + 0.5 * k * (k + 1) * g * dt * dt, rtol=0,
                       atol=1e-13)
    # with an empty active set the step is the unconstrained map's (:66-78): Verlet under
    # constant gravity is the exact parabola
    free = oracle.Sim(dim, q0[dim:], v0[dim:], [rad], [m], np.zeros(dim), up)
    for _ in range(k):
        assert free.step(dt=dt, g=list(g), mu=0.5, integrator=integ, staggering=2)["ncontacts"] == 0
    t = k * dt
    assert np.allclose(free.q, q0[dim:] + v0[dim:] * t + 0.5 * g * t * t, rtol=0, atol=1e-12)
    assert np.allclose(free.v, v0[dim:] + g * t, rtol=0, atol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_empty_active_set_symplectic_euler_closed_form(dim):
    rng = np.random.default_rng(4)
    n = 5
    q0 = rng.uniform(-1, 1, (n, dim)) * np.arange(1, n + 1)[:, None] * 10
    v0 = rng.normal(0, 1, (n, dim))
    g = np.array([0.0, -G, 0.0][:dim])
    dt = 0.02
    sim = oracle.Sim(dim, q0.ravel(), v0.ravel(), np.full(n, 0.01), np.ones(n), np.zeros((0, dim)),
                     np.zeros((0, dim)))
    st = sim.step(dt=dt, g=list(g), mu=0.5, integrator=1, staggering=1)
    assert st["ncontacts"] == 0
    v1 = v0 + g * dt
    assert np.allclose(sim.v.reshape(n, dim), v1, rtol=0, atol=1e-14)
    assert np.allclose(sim.q.reshape(n, dim), q0 + dt * v1, rtol=0, atol=1e-13)


def test_invalid_staggering_options_are_refused():
    sim = oracle.Sim(3, [0.0, 0.1, 0.0], np.zeros(3), [0.1], [1.0], np.zeros(3), [0.0, 1.0, 0.0])
    with pytest.raises(RuntimeError):
        sim.step(dt=0.01, g=[0, -G, 0], mu=0.5, CoR=0.5, staggering=1)  # CoR must be 0
    with pytest.raises(RuntimeError):
        sim.step(dt=0.01, g=[0, -G, 0], mu=0.5, warm_start=3, staggering=2)  # no cache
    with pytest.raises(RuntimeError):
        sim.step(dt=0.01, g=[0, -G, 0], mu=0.5, staggering=3)

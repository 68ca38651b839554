"""Pins the oracle's 3D contact frames, Delassus blocks, multi-contact Gauss-Seidel, FB
residual and restitution by what the mathematics fixes, independently of the oracle's
own formulas:

* a sphere on an incline: rolling a = 5/7 g sin(theta) (the stick impulse needs the
  tangential Delassus term w_t = 1/m + r^2/I = 7/(2m)) when mu >= 2/7 tan(theta), else
  sliding a = g (sin(theta) - mu cos(theta)) with spin-up 5 mu g cos(theta) / (2 r)
  (modelled on assets/ball2d/shell_scripts/sliding_particle_test.py:38-72, in 3D);
* a vertical stack of spheres at rest on a floor carries k m g dt in its k-th contact
  from the top (the one-dimensional stacks of assets/rigidbody3d/tests_serialization);
* sphere-sphere contact sets conserve linear and angular momentum
  (GeometricImpactFrictionMap.cpp:396-410) and, at CoR = 0, do not add kinetic energy;
  every converged contact satisfies the Coulomb conditions, checked from an independent
  numpy assembly of H and M;
* the contact frame is right-handed and orthonormal, s opposes the tangential relative
  velocity, and below |n x u|^2 < 1e-9 it is minus the least-aligned axis projected
  (rigidbody3d/Constraints/SphereSphereConstraint.cpp:273-300, FrictionUtilities.cpp:19-29,
  Constraint.cpp:10-16), restated here in numpy;
* the FB function is the scalar FB of (mu r_n, u_n) when the tangential parts vanish,
  and phi = x + y - z with z the Jordan square root of x o x + y o y in the second-order
  cone (z o z = x o x + y o y, z in the cone);
* restitution: the post-impact normal velocity is CoR times the impact speed
  (Constraint.cpp:18-34), for a sphere on a floor and a head-on pair, 2D and 3D.
"""

import numpy as np
import pytest

import oracle

G = 9.81


def _rot_to(n):
    """A rotation taking e_y to the unit vector n (Rodrigues)."""
    e = np.array([0.0, 1.0, 0.0])
    v = np.cross(e, n)
    c = float(e @ n)
    if np.linalg.norm(v) < 1e-15:
        return np.eye(3) if c > 0 else np.diag([1.0, -1.0, -1.0])
    K = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    return np.eye(3) + K + K @ K / (1.0 + c)


def _incline(theta, azimuth):
    """Unit plane normal at angle theta from the vertical (rotated about y by azimuth) and
    the downhill unit vector."""
    n = np.array([-np.sin(theta), np.cos(theta), 0.0])
    ca, sa = np.cos(azimuth), np.sin(azimuth)
    Ry = np.array([[ca, 0, sa], [0, 1, 0], [-sa, 0, ca]])
    n = Ry @ n
    g = np.array([0.0, -G, 0.0])
    gt = g - (g @ n) * n
    return n, gt / np.linalg.norm(gt)


@pytest.mark.parametrize("theta_deg,mu,azimuth", [(20, 0.5, 0.0), (35, 0.3, 0.7), (20, 0.05, 1.9),
                                                  (35, 0.1, -0.4), (10, 0.0, 2.5)])
def test_sphere_on_incline_rolls_or_slides_as_the_closed_form(theta_deg, mu, azimuth):
    theta = np.radians(theta_deg)
    n, d = _incline(theta, azimuth)
    rad, m, dt = 0.1, 2.0, 0.005
    sim = oracle.Sim(3, rad * n, np.zeros(3), [rad], [m], np.zeros(3), n)
    vs, ws = [], []
    for _ in range(30):
        st = sim.step(dt=dt, g=[0.0, -G, 0.0], mu=mu, tol=1e-26, max_iters=100, integrator=1)
        vs.append(sim.v.copy())
        ws.append(sim.w.copy())
        assert st["ncontacts"] in (0, 1)
    vs, ws = np.array(vs), np.array(ws)
    # from the third step on every step ends with the contact active and solved exactly
    dv = np.diff(vs[2:], axis=0)
    dw = np.diff(ws[2:], axis=0)
    rolling = mu >= 2.0 / 7.0 * np.tan(theta)
    a = 5.0 / 7.0 * G * np.sin(theta) if rolling else G * (np.sin(theta) - mu * np.cos(theta))
    spin = a / rad if rolling else 5.0 * mu * G * np.cos(theta) / (2.0 * rad)
    axis = np.cross(n, d)
    for k in range(dv.shape[0]):
        # the per-step velocity increment along the slope is a dt; nothing normal to it
        assert abs(dv[k] @ d - a * dt) < 1e-12 * max(1.0, a)
        assert np.linalg.norm(dv[k] - (dv[k] @ d) * d) < 1e-12
        # the spin grows about n x d
        assert abs(dw[k] @ axis - spin * dt) < 1e-11 * max(1.0, spin)
        assert np.linalg.norm(dw[k] - (dw[k] @ axis) * axis) < 1e-11
    # after the solve the normal relative velocity vanishes (CoR = 0)
    assert abs(sim.v2() @ n) < 1e-13
    if rolling:
        # the contact point does not slip: v2 + omega2 x (-r n) has no tangential part
        v2 = sim.v2()
        slip = v2 + np.cross(sim.w, -rad * n)
        assert np.linalg.norm(slip - (slip @ n) * n) < 1e-12


def test_sphere_stack_carries_the_weight_above_each_contact():
    nb, rad, m, dt = 8, 0.05, 1.5, 0.005
    q = np.zeros((nb, 3))
    # slightly overlapping, so that every contact is active from the first step on
    q[:, 1] = rad * (1 - 1e-6) + 2 * rad * (1 - 1e-6) * np.arange(nb)
    sim = oracle.Sim(3, q.ravel(), np.zeros(3 * nb), np.full(nb, rad), np.full(nb, m), np.zeros(3),
                     np.array([0.0, 1.0, 0.0]))
    prm = dict(dt=dt, g=[0.0, -G, 0.0], mu=0.4, tol=1e-26, max_iters=200000, integrator=1)
    for _ in range(4):
        st = sim.step(**prm)
    assert st["succeeded"] and st["ncontacts"] == nb
    c = sim.contacts()
    # body pairs (i, i+1) first, then the floor contact of body 0
    assert np.array_equal(c["a"][:nb - 1], np.arange(nb - 1)) and np.array_equal(c["b"][:nb - 1], np.arange(1, nb))
    assert c["b"][-1] == -1 and c["a"][-1] == 0
    w = m * G * dt
    # pair (i, i+1) carries the nb - 1 - i bodies above it; the floor carries all of them
    expect = np.r_[(nb - 1 - np.arange(nb - 1)) * w, nb * w]
    assert np.allclose(c["alpha"], expect, rtol=1e-10, atol=0)
    assert np.abs(c["beta"]).max() < 1e-12 * w
    assert np.abs(sim.v2()).max() < 1e-12


# ----------------------------------------------------------------------------- sphere sets
def _sphere_problem(seed, nb=6, ncont=9):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.3, 0.3, (nb, 3))
    rad = rng.uniform(0.05, 0.15, nb)
    m = rng.uniform(0.5, 2.0, nb)
    inertia = 0.4 * m * rad * rad
    pairs = set()
    while len(pairs) < ncont:
        i, j = sorted(rng.choice(nb, 2, replace=False))
        pairs.add((int(i), int(j)))
    a, b, bases, pa, pb = [], [], [], [], []
    for i, j in sorted(pairs):
        n = x[i] - x[j]
        n /= np.linalg.norm(n)
        p = x[i] + rad[i] / (rad[i] + rad[j]) * (x[j] - x[i])
        s = np.cross(n, rng.normal(size=3))
        s /= np.linalg.norm(s)
        t = np.cross(n, s)
        a.append(i)
        b.append(j)
        bases.append(np.r_[n, s, t])
        pa.append(p - x[i])
        pb.append(p - x[j])
    v0 = rng.normal(size=(nb, 6))
    return dict(x=x, m=m, inertia=inertia, a=np.array(a), b=np.array(b), bases=np.array(bases), pa=np.array(pa),
                pb=np.array(pb), v0=v0)


def _H(P, nb):
    """Independent assembly of the local-frame constraint Jacobian: rows [n; s; t] of each
    contact over the body dofs [lin | ang]; the angular part of row e is arm x e."""
    C = P["a"].size
    H = np.zeros((3 * C, 6 * nb))
    for k in range(C):
        E = P["bases"][k].reshape(3, 3)  # rows n, s, t
        for side, body, arm, sg in ((0, P["a"][k], P["pa"][k], 1.0), (1, P["b"][k], P["pb"][k], -1.0)):
            for row in range(3):
                H[3 * k + row, 6 * body:6 * body + 3] += sg * E[row]
                H[3 * k + row, 6 * body + 3:6 * body + 6] += sg * np.cross(arm, E[row])
    return H


@pytest.mark.parametrize("seed", range(8))
def test_sphere_sets_conserve_momentum_and_do_not_gain_energy(seed):
    P = _sphere_problem(seed)
    nb, C = P["m"].size, P["a"].size
    mu = np.full(C, 0.6)
    r, vout, err, sweeps = oracle.friction_solve(3, P["m"], P["inertia"], P["a"], P["b"], P["bases"], P["pa"], P["pb"],
                                                 mu, None, 0.0, P["v0"], np.zeros(3 * C), max_iters=400000,
                                                 tol=1e-22)
    v0, v1 = P["v0"], vout.reshape(nb, 6)
    m, I, x = P["m"], P["inertia"], P["x"]

    def momenta(v):
        lin = (m[:, None] * v[:, :3]).sum(0)
        ang = (np.cross(x, m[:, None] * v[:, :3]) + I[:, None] * v[:, 3:]).sum(0)
        ke = 0.5 * (m * (v[:, :3] ** 2).sum(1) + I * (v[:, 3:] ** 2).sum(1)).sum()
        return lin, ang, ke

    l0, a0, k0 = momenta(v0)
    l1, a1, k1 = momenta(v1)
    assert np.abs(l1 - l0).max() < 1e-12 * (1 + np.abs(l0).max())
    assert np.abs(a1 - a0).max() < 1e-12 * (1 + np.abs(a0).max())
    assert err < 1e-22
    assert k1 <= k0 * (1 + 1e-12)
    # independent check of the outgoing velocity and of the Coulomb conditions per contact
    H = _H(P, nb)
    Minv = np.diag(np.concatenate([np.r_[np.full(3, 1 / m[i]), np.full(3, 1 / I[i])] for i in range(nb)]))
    assert np.allclose(vout, v0.ravel() + Minv @ H.T @ r, rtol=0, atol=1e-11)
    u = (H @ vout).reshape(C, 3)
    R = r.reshape(C, 3)
    scale = np.abs(R).max() + np.abs(u).max()
    for k in range(C):
        rn, rT, un, uT = R[k, 0], R[k, 1:], u[k, 0], u[k, 1:]
        assert rn >= -1e-10 * scale and un >= -1e-9 * scale
        assert abs(rn * un) < 1e-9 * scale * scale
        assert np.linalg.norm(rT) <= mu[k] * rn + 1e-9 * scale
        if np.linalg.norm(uT) > 1e-6 * scale:
            # sliding: the friction impulse lies on the cone boundary, opposing the slip
            assert np.allclose(rT, -mu[k] * rn * uT / np.linalg.norm(uT), atol=1e-8 * scale)


def test_single_contact_stick_uses_the_sphere_tangential_delassus_term():
    """A sphere against a plane, sticking: r_T = -u_T / w_t with w_t = 1/m + r^2/I."""
    m, rad = 1.7, 0.2
    inertia = 0.4 * m * rad * rad
    n = np.array([0.0, 1.0, 0.0])
    s, t = np.array([1.0, 0.0, 0.0]), np.array([0.0, 0.0, -1.0])  # n x s = t... right-handed: n x s
    t = np.cross(n, s)
    v0 = np.array([[0.3, -2.0, -0.1, 0.5, 0.2, -0.7]])
    arm = -rad * n
    r, vout, err, _ = oracle.friction_solve(3, [m], [inertia], [0], [-1], np.r_[n, s, t][None], arm[None],
                                            np.zeros((1, 3)), [10.0], None, 0.0, v0, np.zeros(3), tol=1e-28)
    u0 = v0[0, :3] + np.cross(v0[0, 3:], arm)
    wt = 1.0 / m + rad * rad / inertia
    assert abs(r[0] - 2.0 * m) < 1e-13  # normal: stops the approach
    assert abs(r[1] + (u0 @ s) / wt) < 1e-13 and abs(r[2] + (u0 @ t) / wt) < 1e-13


# ----------------------------------------------------------------------------- frames
def _orthogonal_vector(n):
    """FrictionUtilities::orthogonalVector restated: the least-aligned axis (ties toward x,
    then y), with its component along n removed, normalised."""
    ax = np.abs(n)
    if ax[0] <= ax[1] and ax[0] <= ax[2]:
        e = np.array([1.0, 0.0, 0.0])
    elif ax[1] <= ax[2]:
        e = np.array([0.0, 1.0, 0.0])
    else:
        e = np.array([0.0, 0.0, 1.0])
    e = e - (n @ e) * n
    return e / np.linalg.norm(e)


def test_contact_frames_are_the_reference_rule():
    rng = np.random.default_rng(3)
    nb = 60
    rad = rng.uniform(0.05, 0.08, nb)
    q = np.zeros((nb, 3))
    # a loose cluster on a floor: every sphere touches its predecessor
    for i in range(nb):
        if i == 0:
            q[i] = [0.0, rad[0], 0.0]
            continue
        while True:
            j = rng.integers(0, i)
            dirn = rng.normal(size=3)
            dirn /= np.linalg.norm(dirn)
            c = q[j] + (rad[i] + rad[j]) * 0.999 * dirn
            if c[1] >= rad[i] * 0.999:
                q[i] = c
                break
    v = rng.normal(scale=0.3, size=(nb, 3))
    w = rng.normal(scale=2.0, size=(nb, 3))
    # a third of the bodies at rest: their contacts among themselves take the
    # orthogonal-vector branch
    rest = rng.random(nb) < 0.35
    v[rest] = 0.0
    w[rest] = 0.0
    m = 1000.0 * 4 / 3 * np.pi * rad ** 3
    sim = oracle.Sim(3, q.ravel(), v.ravel(), rad, m, np.zeros(3), np.array([0.0, 1.0, 0.0]), w=w.ravel())
    sim.step(dt=1e-4, g=[0.0, 0.0, 0.0], mu=0.5, tol=1e-6, max_iters=10, integrator=1)
    c = sim.contacts()
    assert c["a"].size > nb // 2
    branch = 0
    for k in range(c["a"].size):
        E = c["basis"][k]  # rows: n, s, t
        nn, s, t = E
        assert np.abs(E @ E.T - np.eye(3)).max() < 1e-14
        assert abs(np.linalg.det(E) - 1.0) < 1e-14
        assert np.abs(np.cross(nn, s) - t).max() < 1e-14
        i, j = c["a"][k], c["b"][k]
        if j >= 0:
            p = q[i] + rad[i] / (rad[i] + rad[j]) * (q[j] - q[i])
            u = (v[i] + np.cross(w[i], p - q[i])) - (v[j] + np.cross(w[j], p - q[j]))
            assert np.abs(nn - (q[i] - q[j]) / np.linalg.norm(q[i] - q[j])).max() < 1e-15
        else:
            u = v[i] + np.cross(w[i], -rad[i] * nn)
        if np.sum(np.cross(nn, u) ** 2) < 1e-9:
            branch += 1
            assert np.abs(s + _orthogonal_vector(nn)).max() < 1e-15
        else:
            uT = u - (u @ nn) * nn
            assert s @ uT <= 0.0
            assert np.abs(s + uT / np.linalg.norm(uT)).max() < 1e-12
    assert branch > 0


# ----------------------------------------------------------------------------- FB function
def _jordan(a, b):
    return np.r_[a @ b, a[0] * b[1:] + b[0] * a[1:]]


@pytest.mark.parametrize("dim", [2, 3])
def test_fb_reduces_to_the_scalar_fb_without_tangential_parts(dim):
    rng = np.random.default_rng(dim)
    for _ in range(500):
        mu = rng.uniform(0.05, 1.5)
        rn, un = rng.normal(), rng.normal()
        e = oracle.fb_err(dim, mu, np.r_[rn, np.zeros(dim - 1)], np.r_[un, np.zeros(dim - 1)])
        a, b = mu * rn, un
        assert abs(np.sqrt(e) - abs(a + b - np.hypot(a, b))) < 1e-14 * (1 + abs(a) + abs(b))


@pytest.mark.parametrize("dim", [2, 3])
def test_fb_uses_the_jordan_square_root_of_the_cone(dim):
    """phi = x + y - z with x = (mu r_n, r_T), y = (u_n + mu |u_T|, mu u_T) (De Saxce) and
    z the square root in the Jordan algebra of the second-order cone: z o z = x o x + y o y
    with z in the cone."""
    rng = np.random.default_rng(10 + dim)
    for _ in range(500):
        mu = rng.uniform(0.05, 1.5)
        r, u = rng.normal(size=dim), rng.normal(size=dim)
        x = np.r_[mu * r[0], r[1:]]
        y = np.r_[u[0] + mu * np.linalg.norm(u[1:]), mu * u[1:]]
        phi = oracle.fb_phi(dim, mu, r, u)
        z = x + y - phi
        w = _jordan(x, x) + _jordan(y, y)
        assert np.abs(_jordan(z, z) - w).max() < 1e-12 * (1 + np.abs(w).max())
        assert z[0] >= np.linalg.norm(z[1:]) - 1e-12
        assert abs(oracle.fb_err(dim, mu, r, u) - phi @ phi) < 1e-14 * (1 + phi @ phi)


def test_fb_nan_reads_as_infinite():
    assert oracle.fb_err(3, 0.5, np.array([np.nan, 0.0, 0.0]), np.zeros(3)) == np.inf


# ----------------------------------------------------------------------------- restitution
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("cor", [0.0, 0.35, 0.8, 1.0])
def test_restitution_on_a_floor(dim, cor):
    rad, m, V = 0.1, 1.3, 2.0
    q = np.zeros(dim)
    q[1] = rad + 0.5 * V * 0.01  # the predictor penetrates
    v = np.zeros(dim)
    v[1] = -V
    v[0] = 0.4  # tangential motion (frictionless here)
    sim = oracle.Sim(dim, q, v, [rad], [m], np.zeros(dim), np.eye(dim)[1])
    st = sim.step(dt=0.01, g=[0.0, 0.0, 0.0], mu=0.0, CoR=cor, tol=1e-28, integrator=1 if dim == 3 else 0)
    assert st["ncontacts"] == 1
    v2 = sim.v2()
    assert abs(v2[1] - cor * V) < 1e-13
    assert abs(v2[0] - 0.4) < 1e-15


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("cor", [0.0, 0.5, 1.0])
def test_restitution_head_on(dim, cor):
    rad, U = 0.1, 1.5
    q = np.zeros((2, dim))
    q[1, 0] = 2 * rad + 0.001
    v = np.zeros((2, dim))
    v[0, 0], v[1, 0] = U / 2, -U / 2
    m = np.array([1.0, 3.0])
    sim = oracle.Sim(dim, q.ravel(), v.ravel(), [rad, rad], m, np.zeros((0, dim)), np.zeros((0, dim)))
    st = sim.step(dt=0.01, g=[0.0, 0.0, 0.0], mu=0.3, CoR=cor, tol=1e-28, integrator=1 if dim == 3 else 0)
    assert st["ncontacts"] == 1
    v2 = sim.v2().reshape(2, dim)
    # separating speed = CoR x approach speed, momentum kept
    assert abs((v2[1, 0] - v2[0, 0]) - cor * U) < 1e-13
    assert abs(m @ v2[:, 0] - m @ v[:, 0]) < 1e-13

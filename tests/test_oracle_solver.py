"""Pins the oracle's Coulomb local solve, FB residual and block Gauss-Seidel by what the
mathematics fixes: exact contact modes, brute-force enumeration of contact modes on
tiny problems, mu = 0 gives beta = 0 (GeometricImpactFrictionMap.cpp:382-394),
momentum conservation for body-body sets (:396-410), colouring validity, and the
closed-form integrators under constant gravity.
"""

import itertools

import numpy as np
import pytest

import oracle


# ----------------------------------------------------------------------------- FB residual
@pytest.mark.parametrize("dim", [2, 3])
def test_fb_zero_on_exact_modes(dim):
    mu = 0.4
    # separation: r = 0, u_n > 0 (any tangential velocity)
    assert oracle.fb_err(dim, mu, np.zeros(dim), np.r_[0.3, 0.2 * np.ones(dim - 1)]) < 1e-28
    # sticking: u = 0, r strictly inside the cone
    r = np.r_[1.0, 0.1 * np.ones(dim - 1)]
    assert oracle.fb_err(dim, mu, r, np.zeros(dim)) < 1e-28
    # sliding: u_n = 0, r_T = -mu r_n u_T/|u_T|
    uT = np.r_[0.3, -0.4][: dim - 1]
    r = np.r_[2.0, -mu * 2.0 * uT / np.linalg.norm(uT)]
    assert oracle.fb_err(dim, mu, r, np.r_[0.0, uT]) < 1e-26


@pytest.mark.parametrize("dim", [2, 3])
def test_fb_positive_off_solution(dim):
    mu = 0.4
    # penetrating velocity with no impulse
    assert oracle.fb_err(dim, mu, np.zeros(dim), np.r_[-0.5, np.zeros(dim - 1)]) > 1e-3
    # impulse outside the cone while sticking
    assert oracle.fb_err(dim, mu, np.r_[1.0, np.ones(dim - 1)], np.zeros(dim)) > 1e-3
    # friction in the wrong direction while sliding
    uT = np.r_[0.3, -0.4][: dim - 1]
    r = np.r_[2.0, +mu * 2.0 * uT / np.linalg.norm(uT)]
    assert oracle.fb_err(dim, mu, r, np.r_[0.0, uT]) > 1e-3
    # normal impulse with separating velocity (no complementarity)
    assert oracle.fb_err(dim, mu, np.r_[1.0, np.zeros(dim - 1)], np.r_[1.0, np.zeros(dim - 1)]) > 1e-3


# ----------------------------------------------------------------------------- local solve
@pytest.mark.parametrize("dim", [2, 3])
def test_local_solve_is_a_coulomb_solution(dim):
    rng = np.random.default_rng(7)
    for _ in range(2000):
        mu = rng.uniform(0, 1)
        wn, wt = rng.uniform(0.1, 3), rng.uniform(0.1, 3)
        bt = rng.normal(size=dim)
        r = oracle.local_solve(dim, mu, wn, wt, bt)
        u = np.r_[wn * r[0], wt * r[1:]] + bt
        assert oracle.fb_err(dim, mu, r, u) < 1e-24
        assert r[0] >= 0 and np.linalg.norm(r[1:]) <= mu * r[0] * (1 + 1e-12) + 1e-300


# ----------------------------------------------------------------------------- tiny problems
def _problem_2d(seed, ncontacts=3):
    """A random 2D contact problem: nb bodies, body-body and static contacts."""
    rng = np.random.default_rng(seed)
    nb = 3
    m = rng.uniform(0.5, 2.0, nb)
    a, b, basis = [], [], []
    pairs = [(0, 1), (1, 2), (0, -1), (2, -1), (0, 2)]
    for k in range(ncontacts):
        i, j = pairs[k]
        ang = rng.uniform(0, 2 * np.pi)
        n = np.array([np.cos(ang), np.sin(ang)])
        t = np.array([-n[1], n[0]])
        a.append(i)
        b.append(j)
        basis.append(np.stack([n, t], 1))  # columns n, t
    v0 = rng.normal(size=2 * nb)
    return m, np.array(a), np.array(b), np.array(basis), v0


def _assemble(m, a, b, basis, v0):
    nb = m.size
    C = len(a)
    H = np.zeros((2 * C, 2 * nb))
    for k in range(C):
        B = basis[k].T  # rows n^T, t^T
        H[2 * k:2 * k + 2, 2 * a[k]:2 * a[k] + 2] = B
        if b[k] >= 0:
            H[2 * k:2 * k + 2, 2 * b[k]:2 * b[k] + 2] = -B
    Minv = np.diag(np.repeat(1.0 / m, 2))
    W = H @ Minv @ H.T
    bb = H @ v0
    return W, bb, H, Minv


def _enumerate_2d(W, bb, mu):
    """All solutions of the 2D Coulomb NCP by enumeration of contact modes."""
    C = bb.size // 2
    sols = []
    for modes in itertools.product(range(4), repeat=C):  # 0 separate, 1 stick, 2/3 slide +-
        A = np.zeros((2 * C, 2 * C))
        rhs = np.zeros(2 * C)
        for k, md in enumerate(modes):
            rn, rt = 2 * k, 2 * k + 1
            if md == 0:
                A[rn, rn] = 1
                A[rt, rt] = 1
            elif md == 1:
                A[rn] = W[rn]
                rhs[rn] = -bb[rn]
                A[rt] = W[rt]
                rhs[rt] = -bb[rt]
            else:
                s = 1.0 if md == 2 else -1.0
                A[rn] = W[rn]
                rhs[rn] = -bb[rn]
                A[rt, rt] = 1
                A[rt, rn] = s * mu
        try:
            r = np.linalg.solve(A, rhs)
        except np.linalg.LinAlgError:
            continue
        u = W @ r + bb
        ok = True
        for k, md in enumerate(modes):
            rn_, rt_ = r[2 * k], r[2 * k + 1]
            un, ut = u[2 * k], u[2 * k + 1]
            if md == 0:
                ok &= un >= -1e-12
            elif md == 1:
                ok &= rn_ >= -1e-12 and abs(rt_) <= mu * rn_ + 1e-12
            else:
                s = 1.0 if md == 2 else -1.0
                ok &= rn_ >= -1e-12 and s * ut >= -1e-12
        if ok:
            sols.append(r)
    return sols


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("order", [0, 1])
def test_gs_matches_enumerated_ncp_solution(seed, order):
    m, a, b, basis, v0 = _problem_2d(seed)
    mu = 0.5
    W, bb, H, Minv = _assemble(m, a, b, basis, v0)
    sols = _enumerate_2d(W, bb, mu)
    assert sols, "enumeration found no solution"
    C = a.size
    flat = np.stack([basis[k].T.ravel() for k in range(C)])  # column-major [n t] per contact
    r, vout, err, sweeps = oracle.friction_solve(2, m, None, a, b, flat, None, None, np.full(C, mu), None, 0.0, v0,
                                                 np.zeros(2 * C), max_iters=20000, tol=1e-26, order=order)
    assert err < 1e-26
    # the GS solution is one of the enumerated NCP solutions
    d = min(np.max(np.abs(r - s)) for s in sols)
    assert d < 1e-9
    # outgoing velocity v = v0 + M^-1 H^T r
    assert np.allclose(vout, v0 + Minv @ H.T @ r, atol=1e-12)


def test_basis_layout_column_major():
    # bases are column-major [n t]: the first column (entries 0, 1) is the normal
    m, a, b, basis, v0 = _problem_2d(0, ncontacts=1)
    flat = basis[0].T.ravel()  # column-major flattening of [n t]
    assert np.allclose(flat[:2], basis[0][:, 0])


def test_mu_zero_gives_zero_friction():
    m, a, b, basis, v0 = _problem_2d(3)
    C = a.size
    flat = np.stack([basis[k].T.ravel() for k in range(C)])
    r, _, err, _ = oracle.friction_solve(2, m, None, a, b, flat, None, None, np.zeros(C), None, 0.0, v0,
                                         np.zeros(2 * C), tol=1e-24, max_iters=20000)
    assert np.all(r[1::2] == 0.0)


def test_momentum_conserved_for_body_body_contacts():
    rng = np.random.default_rng(11)
    m = rng.uniform(0.5, 2.0, 4)
    a = np.array([0, 1, 2, 0])
    b = np.array([1, 2, 3, 3])
    flat = []
    for k in range(4):
        ang = rng.uniform(0, 2 * np.pi)
        n = np.array([np.cos(ang), np.sin(ang)])
        flat.append(np.r_[n, -n[1], n[0]])
    v0 = rng.normal(size=8)
    r, vout, err, _ = oracle.friction_solve(2, m, None, a, b, np.array(flat), None, None, np.full(4, 0.7), None, 0.0,
                                            v0, np.zeros(8), tol=1e-20, max_iters=50000)
    p0 = (m[:, None] * v0.reshape(4, 2)).sum(0)
    p1 = (m[:, None] * vout.reshape(4, 2)).sum(0)
    assert np.allclose(p0, p1, atol=1e-12)


def test_max_iters_keeps_best_iterate():
    m, a, b, basis, v0 = _problem_2d(5, ncontacts=5)
    C = a.size
    flat = np.stack([basis[k].T.ravel() for k in range(C)])
    r, vout, err, sweeps = oracle.friction_solve(2, m, None, a, b, flat, None, None, np.full(C, 0.5), None, 0.0, v0,
                                                 np.zeros(2 * C), max_iters=3, tol=0.0)
    assert sweeps == 4  # GSIter ends at max_iters + 1 (So-bogus patch :67-70)
    assert err >= 0.0


# ----------------------------------------------------------------------------- colouring
def test_colouring_is_proper_and_deterministic():
    rng = np.random.default_rng(2)
    nb = 500
    a = rng.integers(0, nb, 3000)
    b = rng.integers(-1, nb, 3000)
    b[b == a] = -1
    nc, nr, col = oracle.colour(nb, a, b)
    assert nc > 0 and col.max() == nc - 1
    for c in range(nc):
        sel = col == c
        bodies = np.r_[a[sel], b[sel][b[sel] >= 0]]
        assert np.unique(bodies).size == bodies.size
    nc2, nr2, col2 = oracle.colour(nb, a, b)
    assert np.array_equal(col, col2) and nc == nc2 and nr == nr2


# ----------------------------------------------------------------------------- integrators
@pytest.mark.parametrize("dim,integrator", [(2, 0), (2, 1), (3, 1)])
def test_integrators_closed_form(dim, integrator):
    g = np.array([0.3, -9.81, 0.7][:dim])
    dt = 0.01
    m = np.array([2.0])
    q = np.array([0.1, 0.2, 0.3][:dim])
    v = np.array([1.0, 0.5, -0.25][:dim])
    q0, v0 = q.copy(), v.copy()
    for k in range(1, 201):
        q, v = oracle.umap_flow(dim, integrator, m, g, dt, q, v)
        t = k * dt
        assert np.allclose(v, v0 + g * t, rtol=0, atol=1e-11)
        if integrator == 0:
            # symplectic Euler: q_k = q0 + v0 t + g dt^2 k(k+1)/2
            assert np.allclose(q, q0 + v0 * t + g * dt * dt * k * (k + 1) / 2, atol=1e-11)
        else:
            # Verlet is exact for constant acceleration
            assert np.allclose(q, q0 + v0 * t + 0.5 * g * t * t, atol=1e-11)


# ----------------------------------------------------------------------------- impulse cache
@pytest.mark.parametrize("dim,mode", [(2, 1), (2, 2), (3, 2), (2, 3), (3, 3)])
def test_cached_impulses_warm_start_a_resting_body(dim, mode):
    """A body resting on the floor receives alpha = m g dt every step (the geometric map
    solves on v0 = -g dt); with the previous step's impulse as warm start the solve
    starts at the solution and needs no sweep, while a cold start needs at least one."""
    g = [0.0, -9.81, 0.0][:dim]
    dt, m, rad = 0.01, 2.0, 0.1
    q = np.zeros(dim)
    q[1] = rad
    planes_x, planes_n = np.zeros(dim), np.eye(dim)[1]
    out = {}
    for warm in (0, mode):
        sim = oracle.Sim(dim, q, np.zeros(dim), [rad], [m], planes_x, planes_n)
        stats = [sim.step(dt=dt, g=g, mu=0.5, tol=1e-12, warm_start=warm) for _ in range(4)]
        out[warm] = (stats, sim.contacts()["alpha"])
    for warm, (stats, alpha) in out.items():
        assert alpha.size == 1 and abs(alpha[0] - m * 9.81 * dt) < 1e-12
    assert all(s["sweeps"] >= 1 for s in out[0][0][1:])
    assert all(s["sweeps"] == 0 for s in out[mode][0][2:])


@pytest.mark.parametrize("dim", [2, 3])
def test_a_step_without_caching_leaves_the_cache(dim):
    """The map writes its cache only when it caches: after a cold step (NONE) the next
    caching step still warm-starts from the last cached impulse (no sweep for the resting
    body), while an empty active set empties the cache."""
    g = [0.0, -9.81, 0.0][:dim]
    q = np.zeros(dim)
    q[1] = 0.1
    sim = oracle.Sim(dim, q, np.zeros(dim), [0.1], [2.0], np.zeros(dim), np.eye(dim)[1])
    for _ in range(3):
        sim.step(dt=0.01, g=g, mu=0.5, tol=1e-12, warm_start=2)
    assert sim.step(dt=0.01, g=g, mu=0.5, tol=1e-12, warm_start=0)["sweeps"] >= 1
    assert sim.step(dt=0.01, g=g, mu=0.5, tol=1e-12, warm_start=2)["sweeps"] == 0
    sim.q[1] = 5.0  # lift it off the floor: empty active set
    sim.v[:] = 0.0
    assert sim.step(dt=0.01, g=g, mu=0.5, tol=1e-12, warm_start=0)["ncontacts"] == 0
    sim.q[1] = 0.1
    sim.v[:] = 0.0
    sim.v[1] = -9.81 * 0.01  # resting: v0 = g dt, as after any step on the floor
    assert sim.step(dt=0.01, g=g, mu=0.5, tol=1e-12, warm_start=2)["sweeps"] >= 1  # cache was emptied
    assert sim.step(dt=0.01, g=g, mu=0.5, tol=1e-12, warm_start=2)["sweeps"] == 0


def test_normal_and_friction_cache_is_2d_only():
    sim = oracle.Sim(3, [0, 0.1, 0], np.zeros(3), [0.1], [1.0], np.zeros(3), [0, 1.0, 0])
    with pytest.raises(RuntimeError):
        sim.step(dt=0.01, g=[0, -9.81, 0], mu=0.5, warm_start=1)


def _frame(n, rng):
    n = n / np.linalg.norm(n)
    if n.size == 2:
        return np.c_[n, [-n[1], n[0]]]
    s = np.cross(n, rng.normal(size=3))
    s /= np.linalg.norm(s)
    return np.c_[n, s, np.cross(n, s)]


@pytest.mark.parametrize("dim", [2, 3])
def test_transported_impulse_is_the_rotation_taking_old_to_new_normal(dim):
    """The transported cache rotates the old world impulse with the normal: lengths are
    kept, the old normal lands on the new one (so a purely normal impulse stays purely
    normal), tangential stays tangential, the angle to the normal is kept, zero stays zero,
    and opposite normals still give a proper rotation."""
    rng = np.random.default_rng(dim)
    for _ in range(50):
        n_old = rng.normal(size=dim)
        n_old /= np.linalg.norm(n_old)
        B = _frame(rng.normal(size=dim), rng)
        r_old = rng.normal(size=dim)
        loc = oracle.transport(n_old, r_old, B)
        f = B @ loc
        assert abs(np.linalg.norm(f) - np.linalg.norm(r_old)) < 1e-9
        assert abs(f @ B[:, 0] - r_old @ n_old) < 1e-9
        pure = oracle.transport(n_old, 2.5 * n_old, B)
        assert abs(pure[0] - 2.5) < 1e-9 and np.abs(pure[1:]).max() < 1e-9
        tang = r_old - (r_old @ n_old) * n_old
        assert abs(oracle.transport(n_old, tang, B)[0]) < 1e-9
    assert np.all(oracle.transport(np.zeros(dim), np.zeros(dim), _frame(np.eye(dim)[0], rng)) == 0.0)
    B = _frame(np.eye(dim)[1], rng)
    flip = oracle.transport(-np.eye(dim)[1], np.r_[0.3, -1.0, 0.2][:dim], B)
    assert abs(np.linalg.norm(flip) - np.linalg.norm(np.r_[0.3, -1.0, 0.2][:dim])) < 1e-9
    assert abs(flip[0] - 1.0) < 1e-9  # the old normal component follows the normal

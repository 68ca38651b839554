"""Helpers of the GPU parity tests: device results checked by the FP64 oracle.

check_with_oracle() hands the device's active set (bodies, bases), its impulses and the
start-of-step state to the oracle's stand-alone friction solve with no sweep, so that the
oracle evaluates, in FP64 and with its own assembly of H and M, (1) the FB residual of the
device's impulses (max over contacts of ||phi||^2, the convergence criterion of the step)
and (2) the outgoing velocity v0 + M^-1 H^T r of those impulses.
"""

import numpy as np

import oracle


def arms(dim, q0, r, a, b, basis):
    """Contact arms at q0 (3D): sphere pairs meet at x_a + r_a / (r_a + r_b) (x_b - x_a)
    (rigidbody3d/RigidBody3DSim.cpp:803-809); a plane contact's arm is -r n."""
    C = a.size
    if dim == 2:
        return None, None
    x = q0.reshape(-1, 3)
    pa = np.zeros((C, 3))
    pb = np.zeros((C, 3))
    bb = b >= 0
    ia, ib = a[bb], b[bb]
    p = x[ia] + (r[ia] / (r[ia] + r[ib]))[:, None] * (x[ib] - x[ia])
    pa[bb] = p - x[ia]
    pb[bb] = p - x[ib]
    n = basis[~bb, 0, :]
    pa[~bb] = -r[a[~bb]][:, None] * n
    return pa, pb


def solver_velocities(dim, v, w):
    """Body velocities in the solver layout: [v] (2D) or [v | omega] (3D) per body."""
    if dim == 2:
        return np.ascontiguousarray(v, dtype=np.float64).ravel()
    return np.concatenate([v.reshape(-1, 3), w.reshape(-1, 3)], axis=1).ravel()


def check_with_oracle(dim, state0, contacts, mu, CoR=0.0):
    """state0 = dict(q, v, w, r, m, inertia) at the start of the step; contacts = the
    device's active set and impulses (Context.contacts()); mu = per-contact friction
    coefficients (scalar or array). Returns (oracle residual of the device impulses,
    oracle's v0 + M^-1 H^T r in the solver layout)."""
    a = contacts["a"].astype(np.int32)
    b = contacts["b"].astype(np.int32)
    C = a.size
    basis = contacts["basis"]
    flat = basis.reshape(C, dim * dim)  # rows n, s, t of each contact = column-major [n s t]
    pa, pb = arms(dim, state0["q"], state0["r"], a, b, basis)
    rl = np.concatenate([contacts["alpha"][:, None], contacts["beta"].reshape(C, dim - 1)], axis=1).ravel()
    mu_arr = np.broadcast_to(np.asarray(mu, dtype=np.float64), (C,)).copy()
    v0 = solver_velocities(dim, state0["v"], state0.get("w"))
    ndotv = None
    if CoR != 0.0:
        vv = state0["v"].reshape(-1, dim)
        dv = vv[a] - np.where((b >= 0)[:, None], vv[np.maximum(b, 0)], 0.0)
        ndotv = np.einsum("kd,kd->k", basis[:, 0, :], dv)
    _, vout, err, _ = oracle.friction_solve(dim, state0["m"], state0.get("inertia"), a, b, flat, pa, pb, mu_arr, ndotv,
                                            CoR, v0, rl, max_iters=0, tol=0.0)
    return err, vout


def position_excess(q, q_ref, v_pairs, dt):
    """max |q - q_ref| over what the velocity differences allow: every map's positions are
    q0 plus dt times a combination (weights at most 1) of the start-of-step, predictor and
    solved velocities, so from the same q0 the positions differ by at most
    dt * max |v - v_ref| over the given (v, v_ref) pairs, plus a few units in the last
    place of q. A ratio above 1 is a position update that does not follow the velocities."""
    dq = float(np.abs(np.asarray(q, dtype=np.float64) - np.asarray(q_ref, dtype=np.float64)).max())
    dv = max(float(np.abs(np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64)).max())
             for a, b in v_pairs)
    ulps = 1e-15 * max(1.0, float(np.abs(np.asarray(q_ref)).max()))
    return dq / (dt * dv + ulps)


def rel(x, y, scale=None):
    """max |x - y| over max |y| (or the given scale)."""
    x = np.asarray(x, dtype=np.float64).ravel()
    y = np.asarray(y, dtype=np.float64).ravel()
    if x.size == 0:
        return 0.0
    s = scale if scale is not None else max(1e-300, float(np.abs(y).max()))
    return float(np.abs(x - y).max() / s)


def same_active_set(gc, oc):
    return (np.array_equal(gc["a"], oc["a"]) and np.array_equal(gc["b"], oc["b"]) and
            np.array_equal(gc["plane"], oc["plane"]))

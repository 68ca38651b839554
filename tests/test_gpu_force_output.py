"""The reference's force output (ImpactFrictionMap::exportConstraintForcesToBinaryFile,
scisim/ConstrainedMaps/ImpactFrictionMap.cpp:58-163) from the device, scx_get_contact_forces:
indices (a static plane as -plane - 2), world-space contact points at q0, normals and
impulses alpha n + sum beta_k t_k, checked against the FP64 oracle's active set and
impulses of the same step with the reference's formulas restated in numpy (2D contact
point x_a - r_a n, ball2d/Constraints/BallBallConstraint.cpp:219-222; 3D sphere pairs the
narrow phase's point x_a + r_a / (r_a + r_b) (x_b - x_a), rigidbody3d/RigidBody3DSim.cpp:
803-809; plane contacts x_a - r_a n)."""

import numpy as np
import pytest

import oracle
import scisim_amd as scx
from scisim_amd import scenes

pytestmark = pytest.mark.gpu


def _expected(dim, q0, r, oc):
    a, b, pl, B = oc["a"], oc["b"], oc["plane"], oc["basis"]
    C = a.size
    n = B[:, 0, :]
    x = q0.reshape(-1, dim)
    pts = x[a] - r[a][:, None] * n
    if dim == 3:
        bb = b >= 0
        f = r[a[bb]] / (r[a[bb]] + r[b[bb]])
        pts[bb] = x[a[bb]] + f[:, None] * (x[b[bb]] - x[a[bb]])
    idx = np.stack([a, np.where(b >= 0, b, -pl - 2)])
    frc = oc["alpha"][:, None] * n
    for k in range(1, dim):
        frc = frc + oc["beta"][:, k - 1][:, None] * B[:, k, :]
    return idx, pts.T, n.T, frc.T, C


@pytest.mark.parametrize("dim", [2, 3])
def test_force_output_matches_the_oracle_step(dim):
    if dim == 2:
        s = scenes.ball2d_lattice(m=32, side=5.0, r=0.05, vstd=0.1, seed=2)
        prm = dict(dt=0.01, g=[0.0, -9.81], mu=0.5, CoR=0.0, tol=1e-10, max_iters=5000, integrator=0)
        sim = oracle.Sim(2, s["q"], s["v"], s["r"], s["m"], s["px"], s["pn"])
        sim.run(40, **prm)
    else:
        s = scenes.spheres_pile(n=1500, rmed=0.05, sigma=0.25, seed=9)
        prm = dict(dt=0.005, g=[0.0, -9.81, 0.0], mu=0.4, CoR=0.0, tol=1e-10, max_iters=3000, integrator=1)
        sim = oracle.Sim(3, s["q"], s["v"], s["r"], s["m"], s["px"], s["pn"], inertia=s["inertia"])
    ctx = scx.Context(dim, s["r"].size, max_planes=8)
    try:
        ctx.set_planes(s["px"], s["pn"])
        q0 = sim.q.copy()
        ctx.set_bodies(sim.q, sim.v, sim.r, sim.m, omega=sim.w, inertia=sim.inertia)
        st = ctx.step(scx.step_params(precision=scx.FP64, **prm))
        sim.step(order=1, **prm)
        out = ctx.contact_forces()
        idx, pts, nrm, frc, C = _expected(dim, q0, s["r"], sim.contacts())
        print("dim", dim, "contacts", st["ncontacts"], "planes among them", int((idx[1] < -1).sum()), flush=True)
        assert out["collision_count"] == C == st["ncontacts"] > 0
        assert (idx[1] < -1).any()
        assert np.array_equal(out["collision_indices"], idx)
        assert np.array_equal(out["collision_points"], pts)
        assert np.array_equal(out["collision_normals"], nrm)
        scale = np.abs(frc).max()
        assert np.abs(out["collision_forces"] - frc).max() <= 1e-9 * scale
    finally:
        ctx.close()

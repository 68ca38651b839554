"""The reference-shaped C++ host interface (include/scx_scisim.hpp) and the
FrictionSolver entry point (scx_friction_solve) on the GPU.

* tests/cpp/spatial_grid_tests.cpp restates ball2dtests/collision_detection_tests.cpp
  against the C++ interface and is run on the reference's three AABB fixtures.
* tests/cpp/slope_tests.cpp steps a ball on the reference's inclined-plane scenes
  (assets/ball2d/tests_friction/ball_plane/mu_0p5/*) through GeometricImpactFrictionMap
  and compares with the closed form at t = 0.05 s (5000 steps); the tolerances are the
  reference's (ball2dcli/CMakeLists.txt:74-122), which it registers for t = 3 s.
* scx_friction_solve is compared with the oracle's stand-alone solve on the same
  explicit contact set (2D and 3D), in colour order.
"""

import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
import scisim_amd as scx

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin")


def _run(args, timeout=120):
    print("run", args, flush=True)
    p = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    print(p.stdout, p.stderr, flush=True)
    return p


@pytest.mark.parametrize("name", ["test_00_data", "test_01_data", "test_02_data"])
def test_cpp_spatial_grid(golden_dir, tmp_path, name):
    a = np.load(os.path.join(golden_dir, "aabb_fixtures.npz"))[name]
    path = tmp_path / (name + ".bin")
    a.astype(np.float64).tofile(path)
    p = _run([os.path.join(BIN, "spatial_grid_tests"), str(path)])
    assert p.returncode == 0
    assert "Active sets agree:       1" in p.stdout


@pytest.mark.parametrize("idx", [0, 2, 10, 12, 17])
def test_cpp_slope(golden_dir, idx):
    sc = json.load(open(os.path.join(golden_dir, "slope_scenes.json")))[idx]
    b = sc["ball"]
    args = [os.path.join(BIN, "slope_tests")] + [repr(x) for x in sc["plane_x"] + sc["plane_n"]] + \
        [repr(b[5]), repr(b[4]), repr(sc["dt"]), "5000", repr(sc["mu"])] + [repr(t) for t in sc["tol_x_y_vx_vy"]]
    p = _run(args, timeout=200)
    assert p.returncode == 0


def _random_contact_set(dim, seed, nb=40, C=90):
    rng = np.random.default_rng(seed)
    m = rng.uniform(0.5, 2.0, nb)
    body0 = rng.integers(0, nb, C)
    body1 = rng.integers(-1, nb, C)
    body1[body1 == body0] = -1
    bases, arm0, arm1 = [], [], []
    for k in range(C):
        n = rng.normal(size=dim)
        n /= np.linalg.norm(n)
        if dim == 2:
            bases.append(np.r_[n, -n[1], n[0]])
        else:
            s = np.cross(n, rng.normal(size=3))
            s /= np.linalg.norm(s)
            t = np.cross(n, s)
            bases.append(np.r_[n, s, t])
            la, lb = rng.uniform(0.02, 0.1), rng.uniform(0.02, 0.1)
            arm0.append(-la * n)
            arm1.append(lb * n if body1[k] >= 0 else np.zeros(3))
    v0 = rng.normal(size=dim * nb)
    out = dict(m=m, body0=body0.astype(np.int32), body1=body1.astype(np.int32), bases=np.array(bases),
               mu=rng.uniform(0.1, 0.8, C), v0=v0)
    if dim == 3:
        out["inertia"] = 0.4 * m * rng.uniform(0.02, 0.1, nb) ** 2
        out["omega0"] = rng.normal(size=3 * nb)
        out["arm0"] = np.array(arm0)
        out["arm1"] = np.array(arm1)
    return out


@pytest.mark.parametrize("dim", [2, 3])
def test_friction_solve_matches_oracle(dim):
    p = _random_contact_set(dim, 7 + dim)
    ctx = scx.Context(dim, 64)
    try:
        res = ctx.friction_solve(p["m"], p["v0"], p["body0"], p["body1"], p["bases"], p["mu"],
                                 inertia=p.get("inertia"), omega0=p.get("omega0"), arm0=p.get("arm0"),
                                 arm1=p.get("arm1"), tol=1e-20, max_iters=20000)
    finally:
        ctx.close()
    C = len(p["body0"])
    nb = p["m"].size
    if dim == 2:
        v0s = p["v0"]
    else:
        v0s = np.c_[p["v0"].reshape(nb, 3), p["omega0"].reshape(nb, 3)].ravel()
    r, vout, err, sweeps = oracle.friction_solve(dim, p["m"], p.get("inertia"), p["body0"], p["body1"], p["bases"],
                                                 p.get("arm0"), p.get("arm1"), p["mu"], None, 0.0, v0s,
                                                 np.zeros(dim * C), max_iters=20000, tol=1e-20, order=1)
    print("gpu", res["sweeps"], res["error"], "oracle", sweeps, err, flush=True)
    assert res["succeeded"] and err < 1e-20
    rl = r.reshape(C, dim)
    scale = np.abs(rl).max()
    assert np.abs(res["alpha"] - rl[:, 0]).max() <= 1e-9 * scale
    assert np.abs(res["beta"] - rl[:, 1:]).max() <= 1e-9 * scale
    if dim == 2:
        assert np.allclose(res["v2"], vout, atol=1e-9 * np.abs(vout).max())
    else:
        vv = vout.reshape(nb, 6)
        assert np.allclose(res["v2"].reshape(nb, 3), vv[:, :3], atol=1e-9 * np.abs(vv).max())
        assert np.allclose(res["omega2"].reshape(nb, 3), vv[:, 3:], atol=1e-9 * np.abs(vv).max())
    # friction-only impulse: equals M (v2 - v0) minus the normal impulses' contribution
    mom = np.zeros(dim * nb)
    for k in range(C):
        n = p["bases"][k][:dim]
        for side, i in ((1.0, p["body0"][k]), (-1.0, p["body1"][k])):
            if i >= 0:
                mom[dim * i:dim * i + dim] += side * res["alpha"][k] * n
    lin_f = res["f"][:dim * nb]
    assert np.allclose(lin_f + mom, np.repeat(p["m"], dim) * (res["v2"] - p["v0"]), atol=1e-9)

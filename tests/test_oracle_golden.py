"""Pins the FP64 CPU oracle against the reference's own test data.

* broad phase: ball2dtests/collision_detection_tests.cpp:12-52 (grid == all pairs on
  the three AABB fixtures),
* CCD: scisimtests/narrowphase_tests.cpp:7-455 (11 known-answer tests),
* friction: the analytic ball-on-slope tests registered in ball2dcli/CMakeLists.txt:74-122
  (assets/ball2d/shell_scripts/{sliding,static}_particle_test.*), with the reference's
  tolerances,
* sphere mass/inertia: rigidbody3dtests/rigidbody3d_inertia_tests.cpp:17-56.
"""

import json
import math
import os

import numpy as np
import pytest

import oracle


@pytest.fixture(scope="module")
def aabbs(golden_dir):
    d = np.load(os.path.join(golden_dir, "aabb_fixtures.npz"))
    return {k: d[k] for k in d.files}


@pytest.mark.parametrize("name,expected", [("test_00_data", 94892), ("test_01_data", 637148),
                                           ("test_02_data", 257617)])
def test_grid_equals_all_pairs_on_reference_fixtures(aabbs, name, expected):
    a = aabbs[name]
    mn, mx = a[:, :2], a[:, 2:]
    grid = oracle.grid_pairs(2, mn, mx)
    brute = oracle.all_pairs(2, mn, mx)
    assert np.array_equal(grid, brute)
    assert len(grid) == expected
    # lexicographic, i < j, no duplicates
    assert np.all(grid[:, 0] < grid[:, 1])
    key = grid[:, 0].astype(np.int64) * (1 << 32) + grid[:, 1]
    assert np.all(np.diff(key) > 0)


def test_all_pairs_matches_numpy_brute_force(aabbs):
    a = aabbs["test_00_data"][:1500]
    mn, mx = a[:, :2], a[:, 2:]
    sep = (mx[:, None, :] < mn[None, :, :]).any(-1) | (mx[None, :, :] < mn[:, None, :]).any(-1)
    ii, jj = np.nonzero(np.triu(~sep, 1))
    ref = np.stack([ii, jj], 1).astype(np.uint32)
    assert np.array_equal(oracle.all_pairs(2, mn, mx), ref)


def test_grid_3d_equals_brute_force():
    rng = np.random.default_rng(3)
    n = 3000
    c = rng.uniform(-2, 2, (n, 3))
    h = rng.uniform(0.02, 0.3, (n, 3))
    mn, mx = c - h, c + h
    assert np.array_equal(oracle.grid_pairs(3, mn, mx), oracle.all_pairs(3, mn, mx))


def test_touching_boxes_overlap():
    # AABB::overlaps uses strict max < min: shared faces overlap
    mn = np.array([[0.0, 0.0], [1.0, 0.0], [2.0 + 1e-12, 0.0]])
    mx = np.array([[1.0, 1.0], [2.0, 1.0], [3.0, 1.0]])
    assert oracle.grid_pairs(2, mn, mx).tolist() == [[0, 1]]


@pytest.fixture(scope="module")
def kats(golden_dir):
    return json.load(open(os.path.join(golden_dir, "ccd_kats.json")))


def test_ccd_known_answers(kats):
    assert len(kats) == 11
    for k in kats:
        c = oracle.ccd_coeffs(k["q0a"], k["q1a"], k["ra"], k["q0b"], k["q1b"], k["rb"])
        if k["cexpected"] is not None:
            assert np.all(np.abs(c - np.array(k["cexpected"])) <= 1e-9), k["id"]
        hit, t = oracle.ccd_happens(c)
        assert hit == k["collides"], k["id"]
        if hit:
            assert abs(t - k["t"]) <= k.get("t_tol", 1e-9), (k["id"], t, k["t"])


def _slope_analytic(sc, time_):
    n = np.array(sc["plane_n"])
    n = n / np.linalg.norm(n)
    g, mu = 10.0, 0.5  # hard-coded in the reference's checker, as there
    theta = abs(math.atan2(n[0], n[1]))
    if sc["kind"] == "static":
        return 0.0, 0.0, 0.0, 0.0
    dpl = 0.5 * time_ * time_ * (-g * math.sin(theta) + g * mu * math.cos(theta))
    vpl = time_ * (-g * math.sin(theta) + g * mu * math.cos(theta))
    dx, vx = dpl * math.cos(theta), vpl * math.cos(theta)
    if n[0] > 0:
        dx, vx = -dx, -vx
    return dx, dpl * math.sin(theta), vx, vpl * math.sin(theta)


@pytest.mark.parametrize("idx", list(range(18)))
def test_slope_scenes_match_closed_form(golden_dir, idx):
    sc = json.load(open(os.path.join(golden_dir, "slope_scenes.json")))[idx]
    b = sc["ball"]
    sim = oracle.Sim(2, [b[0], b[1]], [b[2], b[3]], [b[5]], [b[4]], sc["plane_x"], sc["plane_n"])
    nsteps = int(round(sc["end_time"] / sc["dt"]))
    st = sim.run(nsteps, sc["dt"], sc["g"], sc["mu"], CoR=sc["CoR"], tol=sc["tol"], max_iters=sc["max_iters"],
                 eval_every=sc["eval_every"], integrator=1 if sc["integrator"] == "verlet" else 0)
    assert st["failures"] == 0
    exp = _slope_analytic(sc, nsteps * sc["dt"])
    got = (sim.q[0], sim.q[1], sim.v[0], sim.v[1])
    for g_, e_, tol in zip(got, exp, sc["tol_x_y_vx_vy"]):
        assert abs(g_ - e_) <= tol, (sc["name"], got, exp)


def test_sphere_mass_inertia_known_answer():
    # RigidBodySphere::computeMassAndInertia with r = 2.8575, density 1.7404
    r, rho = 2.8575, 1.74040
    M0 = 4.0 * math.pi * r * r * r / 3.0
    I = 2.0 * M0 * r * r / 5.0 * rho
    M = M0 * rho
    assert abs(M - 170.096900946) <= 1e-6
    assert abs(I - 555.557315361) <= 1e-6
    # the product path's default inertia, 2/5 m r^2 from the mass, agrees
    assert abs(2.0 * M * r * r / 5.0 - 555.557315361) <= 1e-6

"""Reference scene files run unchanged on the GPU through the scene-file front end
(include/scx_scene_xml.hpp driven by tests/cpp/scene_xml_tests.cpp, the way
ball2d_cli / rigidbody3d_cli run a scene): the geometric map with the impulse cache and
restitution, the stabilized map in 2D and 3D, and the symplectic-Euler staggering with
stabilization and the transported cache, from the reference's own test assets (copied
under tests/golden/scenes/). After the driver's steps on the device the state must agree
with the FP64 oracle stepping the same parsed scene (colour order) to rounding level."""

import json
import os
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "scene_xml_tests")
SCENES = os.path.join(ROOT, "tests", "golden", "scenes")

CASES = ["balls_falling_on_wedge.xml", "drift_safe_symmetric_break.xml", "drift_safe_balls_on_planes_01.xml",
         "ball_slide_at_threshold_big_dt.xml", "spheres_in_planes_1.xml", "balls_rolling_in_container.xml",
         "rb3d_drift_safe_balls_on_planes_01.xml", "sphere_falling_on_plane.xml", "spheres_in_planes_0.xml"]


@pytest.mark.parametrize("name", CASES)
def test_reference_scene_runs_on_the_device_like_the_oracle(tmp_path, name):
    path = os.path.join(SCENES, name)
    p = subprocess.run([BIN, "parse", path], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    sc = json.loads(p.stdout)
    steps = 20
    out = str(tmp_path / "state.bin")
    print("run", name, flush=True)
    r = subprocess.run([BIN, "run", path, str(steps), out], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    D = sc["dim"]
    n = len(sc["r"])
    got = np.fromfile(out, dtype=np.float64)
    q, v = got[:D * n], got[D * n:2 * D * n]
    prm = sc["params"]
    sim = oracle.Sim(D, sc["q"], sc["v"], sc["r"], sc["m"], sc["plane_x"], sc["plane_n"],
                     w=sc["omega"] if D == 3 else None, inertia=sc["inertia"] if D == 3 else None)
    for _ in range(steps):
        sim.step(dt=sc["dt"], g=sc["g"][:D], mu=sc["mu"], CoR=sc["CoR"], tol=sc["tol"], max_iters=sc["max_iters"],
                 eval_every=sc["eval_every"], integrator=prm["integrator"], order=1, warm_start=prm["warm_start"],
                 staggering=prm["staggering"], stabilize=prm["stabilize"],
                 penetration_threshold=sc["penetration_threshold"])
    scale_q = max(1.0, float(np.abs(sim.q).max()))
    scale_v = max(1e-3, float(np.abs(sim.v).max()))
    dq = float(np.abs(q - sim.q).max()) / scale_q
    dv = float(np.abs(v - sim.v).max()) / scale_v
    print(name, "relative difference after %d steps: q %.3e v %.3e" % (steps, dq, dv), flush=True)
    assert dq < 1e-10 and dv < 1e-8
    if D == 3:
        w = got[2 * D * n:]
        dw = float(np.abs(w - sim.w).max()) / max(1e-3, float(np.abs(sim.w).max()))
        assert dw < 1e-8

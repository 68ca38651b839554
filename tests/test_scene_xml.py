"""CPU checks of the scene-file front end (include/scx_scene_xml.hpp) through its
command-line driver (tests/cpp/scene_xml_tests.cpp, parse mode, host only).

The fixtures under tests/golden/scenes/ are scene files of the reference's own tests
(assets/ball2d/tests_serialization, tests_ccd, tests_friction; assets/rigidbody3d/
tests_symplectic_euler, tests_serialization), copied as data. Each accepted scene is
compared value for value with an independent reading of the same file by Python's
xml.etree following the reference parsers' rules (ball2dutils/Ball2DSceneParser.cpp:
241-369,587-640,1193-1424,1486-1651; rigidbody3dutils/RigidBody3DSceneParser.cpp:
126-546,1152-1210,1742-1965; rigid-body sphere mass and inertia,
rigidbody3d/Geometry/RigidBodySphere.cpp:10-18; rational dt, scisim/Math/Rational.h:204-258).
Scenes with features this pipeline does not run, malformed documents and invalid solver
settings must be refused (exit status 2) with the reason.
"""

import json
import math
import os
import subprocess
import xml.etree.ElementTree as ET
from fractions import Fraction

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "scene_xml_tests")
SCENES = os.path.join(ROOT, "tests", "golden", "scenes")

ACCEPTED = ["ball_ball_ccd_00.xml", "ball_slide_at_threshold_big_dt.xml", "balls_falling_on_wedge.xml",
            "drift_safe_balls_on_planes_01.xml", "drift_safe_symmetric_break.xml",
            "slope_negx_above_threshold_05_degree.xml", "sphere_falling_on_plane.xml", "spheres_in_planes_0.xml",
            "spheres_in_planes_1.xml", "balls_rolling_in_container.xml", "rb3d_drift_safe_balls_on_planes_01.xml"]
REFUSED = {"balls_bouncing_on_wedge_apgd.xml": "impact_operator", "everything_friction.xml": "planar_portal",
           "one_dim_test_0.xml": "fixed", "card_house.xml": "box", "sphere_in_off_center_cylinder.xml": "static_cylinder"}


def parse_cpp(path):
    p = subprocess.run([BIN, "parse", path], capture_output=True, text=True, timeout=60)
    return p.returncode, p.stdout, p.stderr


def _floats(s):
    return [float(x) for x in s.split()]


def _dt(s):
    f = Fraction(s.strip())
    return float(f.numerator) / float(f.denominator)


def expected(path):
    """The reference parsers' reading of an accepted scene, restated with xml.etree."""
    root = ET.parse(path).getroot()
    e = {"dim": 2 if root.tag == "ball2d_scene" else 3}
    D = e["dim"]
    integ = root.find("integrator")
    e["integrator"] = integ.get("type")
    e["dt"] = _dt(integ.get("dt"))
    s = root.find("sobogus_friction_solver")
    e["mu"], e["CoR"], e["tol"] = float(s.get("mu")), float(s.get("CoR")), float(s.get("tol"))
    e["max_iters"], e["eval_every"] = int(s.get("max_iters")), int(s.get("eval_every"))
    e["staggering"] = s.get("staggering")
    px, pn = [], []
    for p in root.findall("static_plane"):
        px += _floats(p.get("x"))
        pn += _floats(p.get("n"))
    e["plane_x"], e["plane_n"] = px, pn
    g = [0.0, 0.0, 0.0]
    q, v, w, r, m, inertia = [], [], [], [], [], []
    if D == 2:
        for gr in root.findall("gravity"):
            g[0] += float(gr.get("fx"))
            g[1] += float(gr.get("fy"))
        for b in root.findall("ball"):
            q += [float(b.get("x")), float(b.get("y"))]
            v += [float(b.get("vx")), float(b.get("vy"))]
            m.append(float(b.get("m")))
            r.append(float(b.get("r")))
    else:
        for gr in root.findall("near_earth_gravity"):
            g = [a + b for a, b in zip(g, _floats(gr.get("f")))]
        geo = [float(x.get("r")) for x in root.findall("geometry")]
        for b in root.findall("rigid_body_with_density"):
            q += _floats(b.get("x"))
            v += _floats(b.get("v"))
            w += _floats(b.get("omega"))
            rad = geo[int(b.get("geo_idx"))]
            rho = float(b.get("rho"))
            M = 4.0 * math.pi * rad * rad * rad / 3.0
            I = 2.0 * M * rad * rad / 5.0
            r.append(rad)
            m.append(M * rho)
            inertia.append(I * rho)
    e.update(q=q, v=v, omega=w, r=r, m=m, inertia=inertia, g=g)
    # step parameters of the C ABI
    stag = {"geometric": 0, "symplectic_euler": 1, "stabilized": 2}[e["staggering"]]
    cache = s.get("cache_impulses", "none") if (D == 2 or stag == 1) else "none"
    warm = {"none": 0, "normal": 2, "normal_and_friction": 1 if D == 2 else 3}[cache]
    e["params"] = {"integrator": 0 if e["integrator"] == "symplectic_euler" else 1, "staggering": stag,
                   "stabilize": int(stag == 1 and s.get("stabilization") == "1"), "warm_start": warm, "precision": 0}
    return e


@pytest.mark.parametrize("name", ACCEPTED)
def test_reference_scene_reads_like_the_reference_parser(name):
    path = os.path.join(SCENES, name)
    rc, out, err = parse_cpp(path)
    assert rc == 0, err
    got = json.loads(out)
    want = expected(path)
    for key in ("dim", "integrator", "mu", "CoR", "tol", "max_iters", "eval_every", "staggering", "params"):
        assert got[key] == want[key], key
    assert got["dt"] == want["dt"]
    for key in ("q", "v", "omega", "r", "m", "inertia", "plane_x", "plane_n", "g"):
        assert np.array_equal(np.array(got[key], dtype=np.float64), np.array(want[key], dtype=np.float64)), key
    assert len(got["r"]) > 0


@pytest.mark.parametrize("name,reason", sorted(REFUSED.items()))
def test_unsupported_reference_scenes_are_refused(name, reason):
    rc, out, err = parse_cpp(os.path.join(SCENES, name))
    assert rc == 2 and reason in err, (rc, err)


BALL = '<ball x="0" y="0" vx="0" vy="0" m="1" r="0.1" fixed="0"/>'
SOLVER = ('<sobogus_friction_solver mu="0.5" CoR="0.0" max_iters="100" tol="1e-9" eval_every="1" '
          'staggering="geometric" cache_impulses="none"/>')


def _doc(solver=SOLVER, body=BALL, integ='<integrator type="verlet" dt="1/100"/>'):
    return "<ball2d_scene>%s%s%s</ball2d_scene>" % (integ, solver, body)


@pytest.mark.parametrize("text,reason", [
    ("<ball2d_scene><integrator type='verlet' dt='0.01'></ball2d_scene>", "does not close"),
    ("<ball2d_scene><integrator type=verlet dt='0.01'/></ball2d_scene>", "not quoted"),
    ("<ball2d_scene><ball x='0' x='1'/></ball2d_scene>", "repeated"),
    ("<ball2d_scene>", "missing end tag"),
    ("<other_scene/>", "unknown scene root"),
    (_doc(solver=""), "sobogus_friction_solver"),
    (_doc(solver=SOLVER.replace('mu="0.5"', 'mu="-1"')), "mu"),
    (_doc(solver=SOLVER.replace('eval_every="1"', 'eval_every="101"')), "eval_every"),
    (_doc(solver=SOLVER.replace('CoR="0.0"', 'CoR="0.5"').replace('"geometric"', '"symplectic_euler"')), "CoR"),
    (_doc(solver=SOLVER.replace('"geometric"', '"stabilized"').replace('"none"', '"normal"')), "stabilized"),
    (_doc(solver=SOLVER.replace('"geometric"', '"symplectic_euler"').replace('CoR="0.0"', 'CoR="0"')),
     "stabilization"),
    (_doc(solver=SOLVER.replace('"geometric"', '"leapfrog"')), "staggering"),
    (_doc(integ='<integrator type="rk4" dt="0.01"/>'), "integrator type"),
    (_doc(integ='<integrator type="verlet" dt="1.5/2"/>'), "dt"),
    (_doc(body=BALL.replace('fixed="0"', 'fixed="1"')), "fixed"),
    (_doc(body='<static_plane x="0 0 0" n="0 1"/>'), "components"),
    (_doc(body='<scripting callback="f"/>'), "scripting"),
])
def test_malformed_and_invalid_scenes_are_refused(tmp_path, text, reason):
    path = tmp_path / "scene.xml"
    path.write_text(text)
    rc, out, err = parse_cpp(str(path))
    assert rc == 2 and reason in err, (rc, err)


def test_comments_declarations_entities_and_rational_dt(tmp_path):
    text = ('<?xml version="1.0"?>\n<!-- a comment with <tags> -->\n<ball2d_scene>\n'
            '  <!-- inner comment -->\n  <camera cx="0" cy="0" scale_factor="1" fps="50" render_at_fps="0" locked="0"/>\n'
            '  <integrator type="symplectic_euler" dt="3/600"/>\n' + SOLVER + '\n'
            '  <end_time t="2.5"></end_time>\n'
            "  <gravity fx='0' fy='-9.81'/><gravity fx='1' fy='0'/>\n  " + BALL + "\n</ball2d_scene>\n")
    path = tmp_path / "scene.xml"
    path.write_text(text)
    rc, out, err = parse_cpp(str(path))
    assert rc == 0, err
    got = json.loads(out)
    assert got["dt"] == 1.0 / 200.0 and got["end_time"] == 2.5
    assert got["g"] == [1.0, -9.81, 0.0]
    assert got["params"]["integrator"] == 0

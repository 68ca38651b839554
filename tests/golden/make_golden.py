"""Writes the golden fixtures under tests/golden/ from the reference's own test data.

Run once in a container where the reference tree is mounted (default
/root/reference); the outputs are committed so that nothing at test time reads the
reference. Inputs only are taken from the reference's test files (data, not code):

* ball2dtests/collision_detection_tests.cpp:12-16 -- three AABB arrays (5k, 10k, 20k
  boxes, 4 doubles each [minx, miny, maxx, maxy], :27-32) -> aabb_fixtures.npz
* scisimtests/narrowphase_tests.cpp:7-455 -- the CCD known-answer inputs and expected
  values -> ccd_kats.json
* assets/ball2d/tests_friction/ball_plane/mu_0p5/*/*.xml with the tolerances that
  ball2dcli/CMakeLists.txt:74-122 registers -> slope_scenes.json
"""

import json
import os
import re
import sys

import numpy as np

REF = os.environ.get("SCISIM_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def aabb_fixtures():
    src = open(os.path.join(REF, "ball2dtests/collision_detection_tests.cpp")).read()
    out = {}
    for name in ("test_00_data", "test_01_data", "test_02_data"):
        m = re.search(r"const scalar %s\[\] = \{([^}]*)\};" % name, src)
        vals = np.array([float(x) for x in m.group(1).split(",")], dtype=np.float64)
        assert vals.size % 4 == 0
        out[name] = vals.reshape(-1, 4)
    np.savez_compressed(os.path.join(HERE, "aabb_fixtures.npz"), **out)
    return {k: v.shape[0] for k, v in out.items()}


def ccd_kats():
    src = open(os.path.join(REF, "scisimtests/narrowphase_tests.cpp")).read()
    cases = []
    for m in re.finditer(r"static int executeCCDTest(\d\d)\(\)\n\{(.*?)\n\}\n", src, re.S):
        body = m.group(2)

        def vec(nm):
            mm = re.search(r"const Vector2s %s\{ ([^,]+), ([^ ]+) \};" % nm, body)
            return [float(mm.group(1)), float(mm.group(2))]

        def sca(nm):
            mm = re.search(r"constexpr scalar %s\{ ([^ ]+) \};" % nm, body)
            return float(mm.group(1)) if mm else None

        case = {
            "id": int(m.group(1)),
            "q0a": vec("q0a"), "q1a": vec("q1a"), "ra": sca("ra"),
            "q0b": vec("q0b"), "q1b": vec("q1b"), "rb": sca("rb"),
        }
        ce = re.search(r"const Vector3s cexpected\{ ([^,]+), ([^,]+), ([^ ]+) \};", body)
        case["cexpected"] = [float(ce.group(i)) for i in (1, 2, 3)] if ce else None
        case["collides"] = ("Collision incorrectly missed" in body)
        # expected first-contact times, written in closed form in the reference test
        if "collision_happens.second != 0.0" in body:
            case["t"] = 0.0
        elif "( 2.0 / 505.0 ) * ( 310.0 - sqrt(4190.0) )" in body:
            case["t"] = (2.0 / 505.0) * (310.0 - np.sqrt(4190.0))
        elif "( 2.0 * ( 550.0 - sqrt( 22766.0 ) ) ) / 1537.0" in body:
            case["t"] = (2.0 * (550.0 - np.sqrt(22766.0))) / 1537.0
        elif "collision_happens.second - 0.5" in body:
            case["t"] = 0.5
        elif "fabs( collision_happens.second ) > 1.0e-9" in body:
            case["t"] = 0.0
        elif sca("collision_time") is not None:
            case["t"] = sca("collision_time")
            case["t_tol"] = 1.0e-7
        cases.append(case)
    json.dump(cases, open(os.path.join(HERE, "ccd_kats.json"), "w"), indent=1)
    return len(cases)


def slope_scenes():
    cm = open(os.path.join(REF, "ball2dcli/CMakeLists.txt")).read()
    tests = []
    for m in re.finditer(
            r"add_test\( (ball2d_(?:static|sliding)_friction_\w+) assets/shell_scripts/(\w+)\.sh "
            r"assets/(tests_friction/ball_plane/mu_0p5/\S+\.xml) ([\d.e-]+) (\d+) ([\d.e-]+) ([\d.e-]+) "
            r"([\d.e-]+) ([\d.e-]+) \)", cm):
        name, script, xml, end_time = m.group(1), m.group(2), m.group(3), float(m.group(4))
        tol = [float(m.group(i)) for i in (6, 7, 8, 9)]
        x = open(os.path.join(REF, "assets/ball2d", xml)).read()
        integ = re.search(r'<integrator type="(\w+)" dt="([^"]+)"/>', x)
        solver = re.search(r"<sobogus_friction_solver ([^/]*)/>", x).group(1)
        sattr = dict(re.findall(r'(\w+)="([^"]*)"', solver))
        grav = re.search(r'<gravity fx="([^"]+)" fy="([^"]+)"/>', x)
        plane = re.search(r'<static_plane x="([^ ]+) ([^"]+)" n="([^ ]+) ([^"]+)"/>', x)
        ball = re.search(r'<ball x="([^"]+)" y="([^"]+)" vx="([^"]+)" vy="([^"]+)" m="([^"]+)" r="([^"]+)"', x)
        tests.append({
            "name": name, "kind": "static" if script.startswith("static") else "sliding",
            "end_time": end_time, "tol_x_y_vx_vy": tol,
            "integrator": integ.group(1), "dt": float(integ.group(2)),
            "mu": float(sattr["mu"]), "CoR": float(sattr["CoR"]), "max_iters": int(sattr["max_iters"]),
            "tol": float(sattr["tol"]), "eval_every": int(sattr["eval_every"]),
            "g": [float(grav.group(1)), float(grav.group(2))],
            "plane_x": [float(plane.group(1)), float(plane.group(2))],
            "plane_n": [float(plane.group(3)), float(plane.group(4))],
            "ball": [float(ball.group(i)) for i in range(1, 7)],
        })
    json.dump(tests, open(os.path.join(HERE, "slope_scenes.json"), "w"), indent=1)
    return len(tests)


if __name__ == "__main__":
    if not os.path.isdir(REF):
        sys.exit("reference tree not found at %s" % REF)
    print("aabb fixtures:", aabb_fixtures())
    print("ccd kats:", ccd_kats())
    print("slope scenes:", slope_scenes())

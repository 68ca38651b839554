"""The launch-width autotuner's search (scisim_amd/csrc/scx_tune.h, host code) on
simulated launches: it settles on the fastest width of a cost bowl, climbs down when
wider is slower, ignores gains under its margin, settles after untimed short launches,
respects the launch limit and starts new classes from the nearest settled class of the
same kind; and the share of the contacts in the sweeps' component phase, on simulated
contact graphs: it holds on a pile whose late colours are small components, climbs to all
contacts on a sparse graph, retries a share that overflowed only now and then, and backs
off from a workgroup with more contacts of one colour than threads
(tests/cpp/tuner_tests.cpp)."""

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "tuner_tests.cpp")
HDR = os.path.join(ROOT, "scisim_amd", "csrc", "scx_tune.h")
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "tuner_tests")


def test_tuner_search_on_simulated_launches():
    if not os.path.exists(BIN) or os.path.getmtime(BIN) < max(os.path.getmtime(SRC), os.path.getmtime(HDR)):
        os.makedirs(os.path.dirname(BIN), exist_ok=True)
        subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.dirname(HDR), SRC, "-o", BIN], check=True)
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    assert "tuner tests passed" in p.stdout

"""Two processes on one GPU. The device's cooperative queues are limited in number
(HSA_QUEUE_TYPE_COOPERATIVE, hsa/hsa.h), so while this test process holds them a second
process's cooperative launches are refused (hipErrorIllegalState); the library then runs
its persistent kernels on one workgroup with ordinary launches. A child process steps the
same scene as this process, while this process's context (which has launched cooperative
kernels) stays alive, and must produce the same state bit for bit (the sweeps do not
depend on the launch width)."""

import os
import subprocess
import sys

import numpy as np
import pytest

import scisim_amd as scx
from scisim_amd import scenes

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys, time
import numpy as np
sys.path.insert(0, sys.argv[1])
import scisim_amd as scx
from scisim_amd import scenes
s = scenes.spheres_pile(n=4000, rmed=0.05, sigma=0.25, seed=3)
ctx = scx.Context(3, s["r"].size, max_planes=8)
ctx.set_planes(s["px"], s["pn"])
ctx.set_bodies(s["q"], s["v"], s["r"], s["m"], inertia=s["inertia"])
p = scx.step_params(dt=0.005, g=[0.0, -9.81, 0.0], mu=0.4, tol=1e-6, max_iters=3000, integrator=1,
                    precision=scx.FP32_POLISH, warm_start=scx.WARM_TRANSPORT)
t0 = time.perf_counter()
for k in range(int(sys.argv[3])):
    st = ctx.step(p)
q, v, w = ctx.get_state()
np.save(sys.argv[2], np.concatenate([q, v, w]))
print("child: contacts", st["ncontacts"], "colours", st["ncolours"], "sweeps", st["sweeps"],
      "%.2f s" % (time.perf_counter() - t0), flush=True)
ctx.close()
"""


def test_second_process_steps_identically_while_this_one_holds_the_cooperative_queue(tmp_path):
    nsteps = 4
    s = scenes.spheres_pile(n=4000, rmed=0.05, sigma=0.25, seed=3)
    ctx = scx.Context(3, s["r"].size, max_planes=8)
    try:
        ctx.set_planes(s["px"], s["pn"])
        ctx.set_bodies(s["q"], s["v"], s["r"], s["m"], inertia=s["inertia"])
        p = scx.step_params(dt=0.005, g=[0.0, -9.81, 0.0], mu=0.4, tol=1e-6, max_iters=3000, integrator=1,
                            precision=scx.FP32_POLISH, warm_start=scx.WARM_TRANSPORT)
        for _ in range(nsteps):
            st = ctx.step(p)
        print("parent: contacts", st["ncontacts"], "colours", st["ncolours"], "sweeps", st["sweeps"], flush=True)
        assert st["ncontacts"] > 2000  # wide enough for multi-workgroup cooperative sweeps
        q, v, w = ctx.get_state()
        out = str(tmp_path / "child.npy")
        r = subprocess.run([sys.executable, "-c", CHILD, ROOT, out, str(nsteps)], capture_output=True, text=True,
                           timeout=240)
        print(r.stdout, r.stderr, flush=True)
        assert r.returncode == 0
        child = np.load(out)
        assert np.array_equal(child, np.concatenate([q, v, w]))
    finally:
        ctx.close()

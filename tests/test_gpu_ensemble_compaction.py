"""Ensembles: the sweeps over compacted lists of the unfrozen scenes' contacts.

Once enough scenes of an ensemble have met the tolerance, the persistent sweep kernel
walks a list of the other scenes' contacts, colour by colour, instead of skipping the
frozen ones one by one. Contacts of one colour share no body, so the order inside a
colour does not change any update: the iterates, the per-scene sweep counts and
residuals and the reported residual must be those of the uncompacted sweeps bit for bit
(SCX_ENSEMBLE_COMPACT=0 switches the compaction off). Checked on one workgroup, on 16
workgroups (the device-wide barrier path) and on the launch width the cost model picks,
with every scene converging and with a sweep limit that stops some scenes (the
best-iterate return of the unconverged scenes).
"""

import sys

import numpy as np
import pytest

import scisim_amd as scx
from scisim_amd import scenes

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def piles():
    return scenes.ensemble_piles(nscenes=96, nmin=100, nmax=2000, seed=3, workers=8)


def _run(monkeypatch, piles, compact, blocks, prm, steps):
    cat, bo, po, mus, _ = piles
    monkeypatch.setenv("SCX_ENSEMBLE_COMPACT", "1" if compact else "0")
    if blocks:
        monkeypatch.setenv("SCX_FORCE_PERSISTENT_BLOCKS", str(blocks))
    else:
        monkeypatch.delenv("SCX_FORCE_PERSISTENT_BLOCKS", raising=False)
    ctx = scx.Context(3, int(bo[-1]), max_planes=int(po[-1]))
    try:
        ctx.set_autotune(False)
        ctx.set_planes(cat["px"], cat["pn"])
        ctx.set_bodies(cat["q"], cat["v"], cat["r"], cat["m"], inertia=cat["inertia"])
        ctx.set_scenes(bo, po, mus)
        out = []
        for k in range(steps):
            st = ctx.step(scx.step_params(iteration=k + 1, **prm))
            q, v, w = ctx.get_state()
            cnt, err, sw = ctx.scene_stats()
            con = ctx.contacts()
            out.append((st, q, v, w, con["alpha"], con["beta"], cnt, err, sw))
        return out
    finally:
        ctx.close()


@pytest.mark.parametrize("blocks", [1, 16, 0])
@pytest.mark.parametrize("limit", ["converged", "sweep_limit"])
def test_compacted_sweeps_match_uncompacted_bit_for_bit(monkeypatch, piles, blocks, limit):
    prm = dict(dt=0.005, g=[0.0, -9.81, 0.0], mu=0.5, CoR=0.0, eval_every=1, integrator=1,
               warm_start=scx.WARM_TRANSPORT, precision=scx.FP32_POLISH)
    prm.update(dict(tol=1e-6, max_iters=20000) if limit == "converged" else dict(tol=1e-11, max_iters=300))
    _say("case blocks", blocks, "limit", limit)
    # the piles are deposited at rest: the first step has nothing to solve, the next ones do
    ref = _run(monkeypatch, piles, False, blocks, prm, 4)
    got = _run(monkeypatch, piles, True, blocks, prm, 4)
    spread = limited = 0
    for k, (a, b) in enumerate(zip(ref, got)):
        st_a, st_b = a[0], b[0]
        sw = a[8].astype(np.int64)
        done = sw != 0xFFFFFFFF
        _say("step", k, "contacts", st_a["ncontacts"], "sweeps", st_a["sweeps"], st_b["sweeps"], "error",
             st_a["error"], st_b["error"], "scenes at tolerance", int(done.sum()), "their sweeps",
             sw[done].min() if done.any() else None, sw[done].max() if done.any() else None)
        for key in ("ncontacts", "ncolours", "sweeps", "sweeps_fp32", "error", "succeeded"):
            assert st_a[key] == st_b[key], key
        for x, y in zip(a[1:], b[1:]):
            assert np.array_equal(x, y)
        # scenes froze at different sweeps (the compaction had something to do)
        spread += int(done.any() and sw[done].min() < st_a["sweeps"])
        limited += int(not st_a["succeeded"] and done.any())
    assert spread >= 2
    if limit == "sweep_limit":
        assert limited >= 2

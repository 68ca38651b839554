"""The device's 2D continuous collision test on the known-answer cases.

tests/golden/ccd_kats.json holds eleven disc pairs (start and end positions, radii) with
the answer of the reference's swept-disc test (ball2d/Constraints/BallBallConstraint.cpp
isActive, a quadratic in the time of contact on [0, 1]): glancing, touching at the start,
moving apart, approaching in time or not. Each pair becomes a two-disc scene stepped once
on the device with dt = 1, no gravity and velocity q1 - q0 under the symplectic Euler
predictor, so that the predictor reaches exactly q1: the step's active set must hold the
pair exactly when the known answer says it collides (and agree with the oracle's).
"""

import json
import os
import sys

import numpy as np
import pytest

import oracle
import scisim_amd as scx

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def kats():
    with open(os.path.join(os.path.dirname(__file__), "golden", "ccd_kats.json")) as f:
        return json.load(f)


def test_device_active_set_follows_the_ccd_known_answers(kats):
    assert len(kats) == 11
    prm = dict(dt=1.0, g=[0.0, 0.0], mu=0.5, CoR=0.0, tol=1e-10, max_iters=1000, eval_every=1, integrator=0)
    ctx = scx.Context(2, 2, max_planes=4)
    try:
        ctx.set_planes(np.zeros((0, 2)), np.zeros((0, 2)))
        for k in kats:
            q0 = np.array(k["q0a"] + k["q0b"], dtype=np.float64)
            v = np.array(k["q1a"], dtype=np.float64) - np.array(k["q0a"], dtype=np.float64)
            v = np.concatenate([v, np.array(k["q1b"], dtype=np.float64) - np.array(k["q0b"], dtype=np.float64)])
            r = np.array([k["ra"], k["rb"]])
            m = np.ones(2)
            ctx.set_bodies(q0, v, r, m)
            st = ctx.step(scx.step_params(**prm))
            sim = oracle.Sim(2, q0, v, r, m, np.zeros((0, 2)), np.zeros((0, 2)))
            so = sim.step(**prm)
            _say("case", k["id"], "collides", k["collides"], "device contacts", st["ncontacts"],
                 "oracle contacts", so["ncontacts"])
            assert st["ncontacts"] == int(bool(k["collides"])), k["id"]
            assert so["ncontacts"] == st["ncontacts"], k["id"]
    finally:
        ctx.close()

"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the inputs still hash to the recorded digest and the oracle still
reproduces the fixture.  GPU: the HIP engine, through the C-ABI, matches the
fixture to relative L2 <= 1e-12 per evolved variable and the clock exactly.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from make_golden import CASES, input_digest, path  # noqa: E402

from cases import rel_l2  # noqa: E402

TOL = 1e-12


def load(name):
    with np.load(path(name), allow_pickle=False) as z:
        return z["w"], z["clock"], int(z["steps"]), str(z["digest"])


@pytest.mark.parametrize("name", sorted(CASES))
def test_inputs_match_recorded_digest(name):
    rp, w, pos = CASES[name][0]()
    _, _, _, digest = load(name)
    assert input_digest(w, pos) == digest, "input generator changed: rerun tests/golden/make_golden.py"


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_fixture(name):
    from oracle.pyoracle import Oracle

    rp, w, pos = CASES[name][0]()
    ref, clock, steps, _ = load(name)
    o = Oracle(rp, w, pos)
    o.step(steps)
    got = o.state()[: ref.shape[0]]
    got_clock = np.array(o.clock())
    o.close()
    assert np.all(rel_l2(got, ref) <= 1e-14)
    assert np.allclose(got_clock, clock, rtol=1e-15, atol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_hip_matches_fixture(name):
    from smaug_amd.engine import Engine

    rp, w, pos = CASES[name][0]()
    ref, clock, steps, _ = load(name)
    print(f"[golden] {name}: {steps} steps on {rp.n}", flush=True)
    e = Engine(rp, w, pos)
    e.step(steps)
    got = e.state()[: ref.shape[0]]
    it, t, qt, dt = e.clock()
    e.close()
    err = rel_l2(got, ref)
    assert np.all(err <= TOL), err
    assert it == int(clock[0])
    assert np.allclose([t, qt, dt], clock[1:], rtol=1e-14, atol=0)

"""The stratified configurations over their own horizons, with their
domain, hydrostatic background, vertical-velocity driver and fixed time
step at resolutions the oracle finishes in minutes: 3D for its 1000 RK4
steps on a 32 x 24 x 24 interior grid (driven from rest, and from a seeded
perturbation), the 2D slice for the 1500 steps of its loss on 64 x 32.  The
HIP engine against the oracle's state at the horizon
(tests/golden/make_horizon.py writes it; in 3D a sample of cells: two whole
interior planes and 1536 random cells; in 2D every interior cell).

Bar: relative L2 per variable <= 1e-12 at the horizon.  The fixture records
the scheme's own sensitivity: the oracle against itself from inputs moved by
one ulp differs by at most 1.2e-14 (3D driven from rest), 2.2e-13 (3D,
seeded perturbation) and 3.7e-14 (2D slice) at the horizon, so the bar is
not below what any change of operation order can hold.  Mixed precision is
held to the north-star's 1e-5 against the same FP64 fixtures.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))

import make_horizon as mh  # noqa: E402

from smaug_amd.params import npert  # noqa: E402
from tests.cases import rel_l2  # noqa: E402

BAR = 1e-12
CASES = sorted(mh.CASES)


def fixture(name):
    path = mh.path(name)
    if not os.path.exists(path):
        pytest.fail(f"missing fixture {path}: run python tests/golden/make_horizon.py")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", CASES)
def test_fixture_matches_its_inputs(name):
    """the fixture was written from the inputs the case builds now"""
    fx = fixture(name)
    build, steps = mh.CASES[name]
    rp, w, pos = build()
    assert str(fx["digest"]) == mh.input_digest(w, pos)
    assert int(fx["steps"]) == steps and tuple(fx["ni"]) == tuple(rp.ni)
    assert np.array_equal(fx["idx"], mh.sample_cells(rp))
    assert fx["w"].shape == (npert(rp.ndim), fx["idx"].size)
    assert np.isfinite(fx["w"]).all()


@pytest.mark.parametrize("name", CASES)
def test_one_ulp_spread_is_below_the_bar(name):
    """the scheme carries a one-ulp input change to well under the bar over
    the horizon on this grid, on the sample and on the whole grid"""
    fx = fixture(name)
    assert fx["spread_full"].max() < BAR / 4, fx["spread_full"]
    assert fx["spread_sample"].max() < BAR / 4, fx["spread_sample"]
    # the driven fields moved away from rest: the comparison is not of zeros
    assert np.abs(fx["w"][1]).max() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_horizon_matches_oracle(name):
    from smaug_amd.engine import Engine

    fx = fixture(name)
    build, steps = mh.CASES[name]
    rp, w, pos = build()
    k = npert(rp.ndim)
    eng = Engine(rp, w, pos)
    eng.step(steps)
    a = eng.state()[:k]
    it, t, qt, dt = eng.clock()
    eng.close()
    assert np.isfinite(a).all()
    err = rel_l2(a[:, fx["idx"]], fx["w"])
    print(f"{name}: rel_l2 after {steps} steps {err.max():.3e} (one-ulp spread "
          f"{fx['spread_sample'].max():.3e})", flush=True)
    assert it == int(fx["clock"][0]) and qt == pytest.approx(float(fx["clock"][2]), rel=1e-15)
    assert np.all(err <= BAR), err


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_mixed_precision_horizon_within_1e5(name):
    """Mixed precision (FP32 stencil shadows, FP64 state, update, RK
    combination and dt) against the FP64 oracle over the same horizon:
    relative L2 per variable <= 1e-5 (measured 3.4e-6 on the perturbed 3D
    case, 2.6e-7 driven from rest, 6.0e-7 on the 2D slice)."""
    from smaug_amd.engine import Engine
    from smaug_amd.params import PREC_MIXED

    fx = fixture(name)
    build, steps = mh.CASES[name]
    rp, w, pos = build()
    rp.precision = PREC_MIXED
    k = npert(rp.ndim)
    eng = Engine(rp, w, pos)
    eng.step(steps)
    a = eng.state()[:k]
    eng.close()
    assert np.isfinite(a).all()
    err = rel_l2(a[:, fx["idx"]], fx["w"])
    print(f"{name}: mixed precision rel_l2 after {steps} steps {err.max():.3e}", flush=True)
    assert np.all(err <= 1e-5), err

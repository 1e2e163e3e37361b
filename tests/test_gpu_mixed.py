"""Mixed precision (SMAUG_PREC_MIXED): the stencils read FP32 shadows of the
evolved fields and FP32 derived fields; the state, background, wave speeds,
CFL step and RK combination stay FP64.  Bar: relative L2 per evolved
variable <= 1e-5 against the FP64 oracle after the stated number of steps.
"""
import numpy as np
import pytest

from smaug_amd import params as P
from smaug_amd.engine import Engine
from tests.cases import compare, ot_case, strat_case

oracle = pytest.importorskip("oracle.pyoracle")

pytestmark = pytest.mark.gpu

TOL = 1e-5


def run(rp, w, pos, nsteps):
    print(f"[mixed] {rp.ndim}D {rp.n} rkon={rp.rkon} x{nsteps}", flush=True)
    e = Engine(rp, w, pos)
    e.step(nsteps)
    got = e.state()
    e.close()
    o = oracle.Oracle(rp, w, pos)
    o.step(nsteps)
    ref = o.state()
    o.close()
    return compare(rp, got, ref)


@pytest.mark.parametrize("rkon", [0, 1])
def test_mixed_stratified_3d(rkon):
    rp, w, pos = strat_case(ni=(14, 10, 12), rkon=rkon)
    rp.precision = P.PREC_MIXED
    err = run(rp, w, pos, 10)
    print("rel L2 per variable:", err)
    assert np.all(err <= TOL), err
    assert np.any(err > 0), "mixed run identical to FP64: shadows not in use"


def test_mixed_stratified_2d():
    rp, w, pos = strat_case(ni=(24, 16), ndim=2, rkon=1)
    rp.precision = P.PREC_MIXED
    err = run(rp, w, pos, 20)
    print("rel L2 per variable:", err)
    assert np.all(err <= TOL), err


def test_mixed_orszag_tang():
    rp, w, pos = ot_case(ni=32, rkon=1)
    rp.precision = P.PREC_MIXED
    err = run(rp, w, pos, 30)
    print("rel L2 per variable:", err)
    assert np.all(err <= TOL), err


@pytest.mark.parametrize("group", ["flux", "hmom", "hb"])
def test_mixed_term_groups(group):
    rp, w, pos = strat_case(ni=(14, 10, 12), rkon=1)
    rp.precision = P.PREC_MIXED
    rp.terms = {"flux": P.T_FLUX, "hmom": P.T_HMOM | P.T_FLUX, "hb": P.T_HB | P.T_FLUX}[group]
    err = run(rp, w, pos, 3)
    assert np.all(err <= TOL), err

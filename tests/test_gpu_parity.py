"""Parity of the HIP engine (through the C-ABI) against the CPU oracle.

Bar: relative L2 per evolved variable <= 1e-12 (FP64).  The HIP kernels
evaluate the same expressions with a different summation order, so the
expected difference is a few ulp per step.
"""
import numpy as np
import pytest

from smaug_amd import params as P
from smaug_amd.engine import Engine
from tests.cases import compare, ot_case, strat_case

oracle = pytest.importorskip("oracle.pyoracle")

pytestmark = pytest.mark.gpu

TOL = 1e-12

GROUPS = {
    "flux": P.T_FLUX,
    "eterm": P.T_ETERM,
    "grav": P.T_GRAV,
    "hrho": P.T_HRHO,
    "he": P.T_HE,
    "hmom": P.T_HMOM,
    "hb": P.T_HB,
    "source": P.T_SOURCE,
    "all": P.TERMS_ALL,
}


def run_pair(rp, w, pos, nsteps):
    eng = Engine(rp, w, pos)
    orc = oracle.Oracle(rp, w, pos)
    eng.step(nsteps)
    orc.step(nsteps)
    a, b = eng.state(), orc.state()
    ca, cb = eng.clock(), orc.clock()
    eng.close()
    orc.close()
    return a, b, ca, cb


@pytest.mark.parametrize("group", list(GROUPS))
def test_ot_one_step_by_group(group):
    rp, w, pos = ot_case(ni=16, rkon=0)
    rp.terms = GROUPS[group]
    a, b, ca, cb = run_pair(rp, w, pos, 1)
    err = compare(rp, a, b)
    print(f"OT group={group} rel_l2={err}", flush=True)
    assert np.all(err <= TOL), err


@pytest.mark.parametrize("group", list(GROUPS))
def test_strat3d_one_step_by_group(group):
    rp, w, pos = strat_case(ndim=3, rkon=0)
    rp.terms = GROUPS[group]
    a, b, ca, cb = run_pair(rp, w, pos, 1)
    err = compare(rp, a, b)
    print(f"strat3d group={group} rel_l2={err}", flush=True)
    assert np.all(err <= TOL), err


def test_grid_matches_oracle():
    rp, w, pos = strat_case(ndim=3)
    eng = Engine(rp, w, pos)
    orc = oracle.Oracle(rp, w, pos)
    p1, d1 = eng.grid()
    p2, d2 = orc.grid()
    assert np.array_equal(p1, p2) and np.array_equal(d1, d2)


@pytest.mark.parametrize("rkon", [0, 1])
def test_ot_multistep(rkon):
    rp, w, pos = ot_case(ni=24, rkon=rkon)
    a, b, ca, cb = run_pair(rp, w, pos, 30)
    err = compare(rp, a, b)
    print(f"OT rkon={rkon} 30 steps rel_l2={err} clock hip={ca} ref={cb}", flush=True)
    assert abs(ca[3] - cb[3]) <= 1e-12 * abs(cb[3])
    assert np.all(err <= TOL), err


@pytest.mark.parametrize("rkon", [0, 1])
def test_strat3d_multistep(rkon):
    rp, w, pos = strat_case(ndim=3, rkon=rkon)
    a, b, ca, cb = run_pair(rp, w, pos, 10)
    err = compare(rp, a, b)
    print(f"strat3d rkon={rkon} 10 steps rel_l2={err}", flush=True)
    assert np.all(err <= TOL), err


@pytest.mark.parametrize("rkon", [0, 1])
def test_strat2d_multistep(rkon):
    rp, w, pos = strat_case(ni=(20, 12), ndim=2, rkon=rkon)
    a, b, ca, cb = run_pair(rp, w, pos, 10)
    err = compare(rp, a, b)
    print(f"strat2d rkon={rkon} 10 steps rel_l2={err}", flush=True)
    assert np.all(err <= TOL), err


def test_cfl_matches_oracle():
    rp, w, pos = ot_case(ni=16)
    eng = Engine(rp, w, pos)
    orc = oracle.Oracle(rp, w, pos)
    c1, dt1 = eng.cfl()
    c2, dt2 = orc.cfl()
    assert np.array_equal(c1, c2), (c1, c2)
    assert dt1 == dt2


@pytest.mark.parametrize("rkon", [0, 1])
def test_cfl_step_restarts_from_dt(rkon):
    """moddton=1 with DT below the CFL limit: every step takes dt = DT
    (src/iosmaug.c:549-569), so t = n*DT exactly; the states match."""
    rp, w, pos = ot_case(ni=16, rkon=rkon)
    orc = oracle.Oracle(rp, w, pos)
    _, dt_cfl = orc.cfl()
    orc.close()
    rp.dt = dt_cfl / 7.0
    a, b, ca, cb = run_pair(rp, w, pos, 4)
    print(f"OT rkon={rkon} DT below CFL: clock hip={ca} ref={cb}", flush=True)
    assert ca[3] == rp.dt and cb[3] == rp.dt
    assert ca[1] == cb[1] and ca[2] == cb[2]
    err = compare(rp, a, b)
    assert np.all(err <= TOL), err


def test_strat3d_many_interior_tiles():
    """A grid large enough that most workgroups take the interior fast
    paths (no halo substitutions, no boundary zeroing) and the tile order
    spreads bricks of tiles over the XCDs: 32 x 32 x 64 interior, RK4."""
    rp, w, pos = strat_case(ni=(32, 32, 64), ndim=3, rkon=1)
    a, b, ca, cb = run_pair(rp, w, pos, 2)
    err = compare(rp, a, b)
    print(f"strat3d 32x32x64 RK4 x2 rel_l2={err}", flush=True)
    assert np.all(err <= TOL), err


def test_strat2d_many_interior_tiles():
    rp, w, pos = strat_case(ni=(96, 64), ndim=2, rkon=1)
    a, b, ca, cb = run_pair(rp, w, pos, 3)
    err = compare(rp, a, b)
    print(f"strat2d 96x64 RK4 x3 rel_l2={err}", flush=True)
    assert np.all(err <= TOL), err


def test_nonseparable_grid():
    """Positions that are not a tensor product (x stretched with y), so the
    kernels read the full per-cell spacing arrays instead of per-axis ones."""
    rp, w, pos = ot_case(ni=20, rkon=1)
    rp.moddton = 0
    rp.dt = 1.0e-3
    pos = pos.copy()
    pos[0] = pos[0] * (1.0 + 0.02 * np.sin(2 * np.pi * pos[1]))
    a, b, ca, cb = run_pair(rp, w, pos, 3)
    err = compare(rp, a, b)
    print(f"OT non-separable grid RK4 x3 rel_l2={err}", flush=True)
    assert np.all(err <= TOL), err
    eng = Engine(rp, w, pos)
    _, d1 = eng.grid()
    eng.close()
    assert not np.allclose(d1[0].reshape(rp.n[1], rp.n[0])[3], d1[0].reshape(rp.n[1], rp.n[0])[7])


def test_nonseparable_grid_driver():
    """Stratified 3D with the vertical driver on positions that are not a
    tensor product (dim 1 stretched with height): the stage kernel forms the
    driver shape per cell instead of from per-axis factors."""
    rp, w, pos = strat_case(ni=(14, 10, 12), ndim=3, rkon=1)
    pos = pos.copy()
    z = pos[0]
    pos[1] = pos[1] * (1.0 + 0.02 * np.sin(2 * np.pi * (z - z.min()) / (z.max() - z.min())))
    a, b, ca, cb = run_pair(rp, w, pos, 2)
    err = compare(rp, a, b)
    print(f"strat3d non-separable grid with driver RK4 x2 rel_l2={err}", flush=True)
    assert np.all(err <= TOL), err
    eng = Engine(rp, w, pos)
    p1, _ = eng.grid()
    eng.close()
    col = p1[1].reshape(rp.n[2], rp.n[1], rp.n[0])[4, 5]
    assert np.ptp(col) > 0  # position along dim 1 varies along dim 0

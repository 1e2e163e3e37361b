"""The reference's user-source drivers beyond the vertical one: fluxtube1's
torsional driver (models/usersource.cu.fluxtube1:3-90) and spicule1's
(models/usersource.cu.spicule1:4-136).

CPU: the oracle's restatement against properties the reference's formulas
fix without restating them -- the torsional velocity is azimuthal about the
axis through (xc2, xc3) with the kinetic energy of the injected momentum
added to e, vanishes on the axis and peaks at r = width/sqrt(2); the spicule
driver acts in the step whose counter is 1 only, at the grid positions
xmin + i dx (not the cell positions), with the amplitude AA tdep at the
grid node on its centre and no energy term.
GPU: the HIP engine against the oracle, all terms on, Euler and RK4.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle.pyoracle import Oracle
from smaug_amd import configs
from smaug_amd.params import T_SOURCE
from tests.cases import compare, strat_case


def torsional_case(rkon=1, terms=None, ni=(14, 10, 12)):
    rp, w, pos = strat_case(ni=ni, ndim=3, rkon=rkon)
    dx = rp.dx
    configs.fluxtube1_driver(rp, aa=200.0, tdep=1.0, centre=(rp.xmin[0] + 2.5 * dx[0], 2.0e6, 2.0e6),
                             widths=(3.0 * dx[0], 3.0 * dx[1], 3.0 * dx[1]), scales=(2.0e6, 2.0e6))
    if terms is not None:
        rp.terms = terms
    return rp, w, pos


def spicule_case(ndim=2, rkon=0, terms=None, ni=(16, 12, 10)):
    rp, w, pos = strat_case(ni=ni[:ndim], ndim=ndim, rkon=rkon)
    dx = rp.dx
    # centre on the grid node of indices (5, 7): xmin + i dx with i counting ghost cells
    configs.spicule1_driver(rp, AA=200.0, tdep=1.0, centre=(rp.xmin[0] + 5 * dx[0], rp.xmin[1] + 7 * dx[1]),
                            widths=(3.0 * dx[0], 3.0 * dx[1]))
    rp.it0 = 0
    if terms is not None:
        rp.terms = terms
    return rp, w, pos


def oracle_steps(rp, w, pos, n):
    o = Oracle(rp, w, pos)
    out = []
    for _ in range(n):
        o.step(1)
        out.append(o.state().copy())
    o.close()
    return out


def test_torsional_velocity_is_azimuthal_with_its_kinetic_energy():
    rp, w, pos = torsional_case(rkon=0, terms=T_SOURCE)
    (s1,) = oracle_steps(rp, w, pos, 1)
    nd = 3
    dt = rp.dt
    rt = w[0] + w[3 + 2 * nd]
    dm = (s1[1:4] - w[1:4]) / (dt * rt)  # injected velocity per component
    interior = interior_cells(rp)
    # the oracle's whole-grid positions, as smaug_get_grid reports them
    o = Oracle(rp, w, pos)
    gpos, _ = o.grid()
    o.close()
    xp = gpos[1] - rp.drv[3]
    yp = gpos[2] - rp.drv[4]
    v1, v2 = dm[1][interior], dm[2][interior]
    assert np.abs(dm[0][interior]).max() == 0.0  # m_0 untouched
    assert np.abs(v1).max() > 1.0  # the driver is not a no-op here
    # azimuthal: v . r = 0 (xxmax = yymax), to roundoff of the products
    radial = v1 * xp[interior] + v2 * yp[interior]
    scale = np.abs(v1 * xp[interior]).max()
    assert np.abs(radial).max() <= 1e-12 * scale
    # energy = kinetic energy of the injected momentum (rho_t unchanged)
    de = s1[1 + nd] - w[1 + nd]
    ke = 0.5 * dt * rt * (dm[1] ** 2 + dm[2] ** 2)
    assert np.allclose(de[interior], ke[interior], rtol=1e-12, atol=1e-12 * np.abs(ke).max())
    # density and field untouched
    assert np.array_equal(s1[0][interior], w[0][interior])
    assert np.array_equal(s1[2 + nd:2 + 2 * nd][:, interior], w[2 + nd:2 + 2 * nd][:, interior])


def test_torsional_speed_profile_on_a_fine_line():
    """|v| along dim 1 through the axis at the centre height: r exp(-r^2/w^2)
    times the z factor, so it vanishes on the axis and peaks at w/sqrt(2)."""
    rp, w, pos = torsional_case(rkon=0, terms=T_SOURCE, ni=(8, 64, 8))
    width = 8.0 * rp.dx[1]
    rp.drv[6] = rp.drv[7] = width
    (s1,) = oracle_steps(rp, w, pos, 1)
    rt = w[0] + w[3 + 2 * 3]
    v2 = ((s1[3] - w[3]) / (rp.dt * rt)).reshape(rp.n[2], rp.n[1], rp.n[0])
    o = Oracle(rp, w, pos)
    gpos, _ = o.grid()
    o.close()
    g = [p.reshape(rp.n[2], rp.n[1], rp.n[0]) for p in gpos]
    k = int(np.argmin(np.abs(g[2][:, 0, 0] - rp.drv[4])))
    i = int(np.argmin(np.abs(g[0][0, 0, :] - rp.drv[2])))
    line = np.abs(v2[k, 2:-2, i])
    r = np.abs(g[1][k, 2:-2, i] - rp.drv[3])
    # same line evaluated from the closed form, up to the constant factor
    model = r * np.exp(-r * r / width ** 2)
    c = line.max() / model.max()
    assert np.allclose(line, c * model, rtol=1e-9, atol=1e-12 * line.max())
    rpk = r[np.argmax(line)]
    assert abs(rpk - width / np.sqrt(2.0)) <= rp.dx[1]


def interior_cells(rp):
    n = rp.n
    idx = np.indices(tuple(n[:rp.ndim][::-1])).reshape(rp.ndim, -1)[::-1]
    m = np.ones(rp.ncells, bool)
    for d in range(rp.ndim):
        m &= (idx[d] >= 2) & (idx[d] <= n[d] - 3)
    return m


@pytest.mark.parametrize("case", ["vertical", "torsional", "spicule"])
def test_oracle_driver_param_setter_equals_a_fresh_oracle(case):
    """Oracle.set_driver_param (used to difference the loss in the driver
    amplitude) acts as an oracle created with that parameter: the same
    states bit for bit after two steps; the amplitude moves the state."""
    if case == "vertical":
        rp, w, pos = strat_case(ni=(16, 12), ndim=2, rkon=1)
    elif case == "torsional":
        rp, w, pos = torsional_case(rkon=1)
    else:
        rp, w, pos = spicule_case(ndim=2, rkon=1)
    a0 = rp.drv[0]
    o = Oracle(rp, w, pos)
    o.set_driver_param(0, 1.25 * a0)
    o.step(2)
    got = o.state()
    o.close()
    rp.drv = [1.25 * a0] + list(rp.drv[1:])
    o = Oracle(rp, w, pos)
    o.step(2)
    want = o.state()
    o.close()
    rp.drv = [a0] + list(rp.drv[1:])
    o = Oracle(rp, w, pos)
    o.step(2)
    base = o.state()
    o.close()
    assert np.array_equal(got, want)
    assert not np.array_equal(got, base)
    with pytest.raises(ValueError):
        Oracle(rp, w, pos).set_driver_param(99, 1.0)


def test_spicule_acts_in_the_step_with_counter_one_only():
    rp, w, pos = spicule_case(ndim=2, rkon=0, terms=T_SOURCE)
    s = oracle_steps(rp, w, pos, 3)  # steps with counters 0, 1, 2
    m = interior_cells(rp)
    assert np.array_equal(s[0][:6, m], w[:6, m])
    assert not np.array_equal(s[1][1, m], s[0][1, m])
    assert np.array_equal(s[2][:6, m], s[1][:6, m])
    # only m_0 moves; no energy term
    assert np.array_equal(s[1][0, m], s[0][0, m]) and np.array_equal(s[1][2:6, m], s[0][2:6, m])


def test_spicule_uses_grid_node_positions():
    """The centre sits on the grid node of indices (5, 7), xmin + i dx; there
    exp(0) = 1, so dm_0 / (dt rho_t) = AA tdep exactly, and the cell
    positions (cell centres, half a cell away) are not what it reads."""
    rp, w, pos = spicule_case(ndim=2, rkon=0, terms=T_SOURCE)
    s = oracle_steps(rp, w, pos, 2)
    rt = w[0] + w[3 + 2 * 2]
    v = ((s[1][1] - s[0][1]) / (rp.dt * rt)).reshape(rp.n[1], rp.n[0])
    j, i = np.unravel_index(np.argmax(v), v.shape)
    assert (i, j) == (5, 7)
    assert abs(v[7, 5] - 200.0) <= 1e-9 * 200.0
    # one cell away along dim 0 the factor is exp(-dx^2 / (3 dx)^2)
    assert abs(v[7, 6] / v[7, 5] - np.exp(-1.0 / 9.0)) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("rkon", [0, 1])
def test_gpu_torsional_matches_oracle(rkon):
    from smaug_amd.engine import Engine

    rp, w, pos = torsional_case(rkon=rkon)
    eng = Engine(rp, w, pos)
    eng.step(3)
    a = eng.state()
    eng.close()
    b = oracle_steps(rp, w, pos, 3)[-1]
    err = compare(rp, a, b)
    print(f"torsional rkon={rkon} rel_l2={err}", flush=True)
    assert np.all(err <= 1e-12), err


@pytest.mark.gpu
@pytest.mark.parametrize("ndim,rkon", [(2, 0), (2, 1), (3, 1)])
def test_gpu_spicule_matches_oracle(ndim, rkon):
    from smaug_amd.engine import Engine

    rp, w, pos = spicule_case(ndim=ndim, rkon=rkon)
    eng = Engine(rp, w, pos)
    eng.step(3)
    a = eng.state()
    eng.close()
    b = oracle_steps(rp, w, pos, 3)[-1]
    assert not np.array_equal(b[1], w[1])
    err = compare(rp, a, b)
    print(f"spicule {ndim}D rkon={rkon} rel_l2={err}", flush=True)
    assert np.all(err <= 1e-12), err


@pytest.mark.gpu
@pytest.mark.parametrize("driver", ["torsional", "spicule"])
def test_gpu_drivers_in_mixed_precision(driver):
    """The drivers act on the FP64 update in mixed precision too: the mixed
    engine against the FP64 oracle, relative L2 <= 1e-5 after 3 RK4 steps."""
    from smaug_amd.engine import Engine
    from smaug_amd.params import PREC_FP64, PREC_MIXED

    rp, w, pos = torsional_case(rkon=1) if driver == "torsional" else spicule_case(ndim=2, rkon=1)
    rp.precision = PREC_MIXED
    eng = Engine(rp, w, pos)
    eng.step(3)
    a = eng.state()
    eng.close()
    rp.precision = PREC_FP64
    b = oracle_steps(rp, w, pos, 3)[-1]
    err = compare(rp, a, b)
    print(f"{driver} mixed rel_l2={err}", flush=True)
    assert np.all(err <= 1e-5), err

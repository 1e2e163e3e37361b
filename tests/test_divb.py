"""divB source terms (divbon=1, divbfix=1; cudivb -> divb_parallel,
src/divb.cu:165-338, called after cugrav at src/iosmaug.c:678-679): the
oracle pinned by closed forms on the CPU, the HIP engine against the oracle
on the GPU.

divb = d1 b1 + d2 b2 + d1 b1b + d2 b2b (grad3d: 4th order over 2 p->dx,
exact on polynomials of degree <= 4), on cells 2 < i,j < n-2 with the
layers 3, n-4, n-3 zeroed; m_d += dt divb b_d, b_d += dt divb m_d/rho_t.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle.pyoracle import Oracle
from smaug_amd import configs
from smaug_amd.params import BC_CONTCD4, RunParams, nvar

from cases import compare, strat_case


def divb_case(ndim=2, divbfix=1, ni=(14, 12, 10)):
    rp = RunParams(ndim=ndim, ni=list(ni[:ndim]), xmin=[0.0] * ndim, xmax=[1.0] * ndim, dt=1e-3,
                   gamma=5.0 / 3.0, courant=0.3, rkon=0, moddton=0, hyperdifmom=1, chyp={"default": 0.02},
                   bcu=[BC_CONTCD4] * 3, bcl=[BC_CONTCD4] * 3, terms=0, divbon=1, divbfix=divbfix)
    pos = np.stack(configs._mesh(rp))
    nd = ndim
    w = np.zeros((nvar(nd), rp.ncells))
    w[2 + 2 * nd] = 0.6 / (rp.gamma - 1.0)
    w[3 + 2 * nd] = 1.0
    return rp, w, pos


def coords(rp):
    n = rp.n
    idx = np.arange(rp.ncells)
    return idx % n[0], (idx // n[0]) % n[1]


def expected_divb(rp, gx, gy):
    """divb per cell for constant gradients gx = d1(b1 + b1b), gy = d2(b2 + b2b):
    nonzero on 2 < i,j < n-2, each term zeroed on its layers 3, n-4, n-3."""
    i, j = coords(rp)
    n0, n1 = rp.n[0], rp.n[1]
    region = (i > 2) & (j > 2) & (i < n0 - 2) & (j < n1 - 2)
    zx = (i == 3) | (i == n0 - 4) | (i == n0 - 3)
    zy = (j == 3) | (j == n1 - 4) | (j == n1 - 3)
    return np.where(region, np.where(zx, 0.0, gx) + np.where(zy, 0.0, gy), 0.0), region


def run_oracle(rp, w, pos, n=1):
    o = Oracle(rp, w, pos)
    o.step(n)
    out = o.state()
    o.close()
    return out


def test_divb_sources_closed_form():
    """Linear b (perturbation and background), uniform momentum: divb is the
    constant sum of the slopes where not zeroed, and the momentum and field
    updates are dt divb b_d and dt divb m_d/rho_t; rho and e do not change."""
    rp, w, pos = divb_case()
    x, y = pos
    a, c, ab, cb = 0.3, -0.7, 0.11, 0.05
    w[4] = 0.2 + a * x          # b1
    w[5] = -0.1 + c * y         # b2
    w[8] = 0.4 + ab * x         # b1b
    w[9] = 0.3 + cb * y         # b2b
    w[1], w[2] = 0.02, -0.03    # m
    w[0] = 0.5                  # rho_t = 1.5
    out = run_oracle(rp, w, pos)
    dv, region = expected_divb(rp, a + ab, c + cb)
    i, j = coords(rp)
    n0, n1 = rp.n[0], rp.n[1]
    interior = (i >= 2) & (j >= 2) & (i <= n0 - 3) & (j <= n1 - 3)
    dt = rp.dt
    for d, (fm, fb) in enumerate(((1, 4), (2, 5))):
        assert np.allclose(out[fm][interior], (w[fm] + dt * dv * w[fb])[interior], rtol=1e-13, atol=1e-16), d
        want = w[fb] + dt * dv * w[fm] / 1.5
        assert np.allclose(out[fb][interior], want[interior], rtol=1e-13, atol=1e-16), d
    assert np.array_equal(out[0][interior], w[0][interior])
    assert np.array_equal(out[3][interior], w[3][interior])
    assert np.any(dv != 0) and np.any(region & (dv == 0))


def test_divb_vanishes_for_divergence_free_fields():
    """b1 = f(y), b2 = g(x): every difference across the field's own axis
    is of a constant, so divb and the update are exactly zero."""
    rp, w, pos = divb_case()
    x, y = pos
    w[4] = np.sin(2 * np.pi * y)
    w[5] = np.cos(2 * np.pi * x)
    w[1], w[2] = 0.02, -0.03
    out = run_oracle(rp, w, pos)
    i, j = coords(rp)
    n0, n1 = rp.n[0], rp.n[1]
    interior = (i >= 2) & (j >= 2) & (i <= n0 - 3) & (j <= n1 - 3)
    assert np.array_equal(out[:6, interior], w[:6, interior])


@pytest.mark.parametrize("ndim,divbfix", [(2, 0), (3, 1)])
def test_divb_inactive_without_fix_or_in_3d(ndim, divbfix):
    """divbfix = 0 leaves only the zeroing of the temporary; in 3D the
    reference's launch covers the k = 0 plane only, which its region test
    excludes, so nothing changes either."""
    rp, w, pos = divb_case(ndim=ndim, divbfix=divbfix)
    x = pos[0]
    w[2 + ndim] = 0.2 + 0.3 * x  # b1, divergent
    w[1] = 0.02
    out = run_oracle(rp, w, pos)
    k = 2 + 2 * ndim
    i = np.arange(rp.ncells) % rp.n[0]
    inner = (i >= 2) & (i <= rp.n[0] - 3)
    assert np.array_equal(out[:k, inner], w[:k, inner]) or ndim == 3 and np.allclose(out[:k], w[:k])


@pytest.mark.gpu
@pytest.mark.parametrize("rkon", [0, 1])
def test_gpu_divb_matches_oracle(rkon):
    """HIP (through the C-ABI) against the oracle with divbon=1, divbfix=1
    on a 2D stratified state with every other term on."""
    from smaug_amd.engine import Engine

    rp, w, pos = strat_case(ni=(20, 16), ndim=2, rkon=rkon)
    rp.divbon, rp.divbfix = 1, 1
    w = w.copy()
    z = pos[0]
    w[4] += 1e-3 * np.sin(2 * np.pi * (z - z.min()) / (z.max() - z.min())) * w[8]  # divergent b1
    eng = Engine(rp, w, pos)
    eng.step(3)
    a = eng.state()
    eng.close()
    b = run_oracle(rp, w, pos, n=3)
    # the divB terms are not negligible in this case
    rp0 = strat_case(ni=(20, 16), ndim=2, rkon=rkon)[0]
    c = run_oracle(rp0, w, pos, n=3)
    assert np.max(compare(rp, b, c)) > 1e-10
    err = compare(rp, a, b)
    print(f"divb rkon={rkon} rel_l2={err}", flush=True)
    assert np.all(err <= 1e-12), err

"""Adjoint of the time loop (smaug_adjoint_step behind smaug_amd.adjoint)
against central finite differences of the oracle with its hyperdiffusion
coefficients replayed from the unperturbed run, i.e. the derivative of the
step with those (nonsmooth, windowed-maximum) coefficients frozen.

Bar: directional derivatives agree to 1e-6 relative, for random directions
in the driver amplitude map and in the initial state, step h = 1e-6 relative.
"""
from __future__ import annotations

import numpy as np
import pytest

from smaug_amd import params as P
from tests.adjoint_cases import ReplayFD, interior_mask
from tests.cases import ot_case, strat_case

torch = pytest.importorskip("torch")
oracle = pytest.importorskip("oracle.pyoracle")

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="the adjoint path's first MI355X run ended in a GPU fault whose cause is not "
                            "identified yet; until it is, the path is gated off (SMAUG_ADJOINT) and checked in "
                            "the host build instead (tests/test_hostemu.py)"),
]

TOL = 1e-6
EPS = 1e-6


def gpu_grads(rp, w0, pos, amap, nsteps, i_line, every):
    import os

    os.environ["SMAUG_ADJOINT"] = "1"
    from smaug_amd.adjoint import driver_loss, line_vz2_loss
    from smaug_amd.engine import Engine

    k = P.npert(rp.ndim)
    eng = Engine(rp, w0, pos)
    dev = torch.device("cuda:0")
    w0t = torch.tensor(np.ascontiguousarray(w0[:k]).ravel(), dtype=torch.float64, device=dev, requires_grad=True)
    ncol = rp.ncells // rp.n[0]
    a = np.zeros(ncol) if amap is None else amap
    amt = torch.tensor(a, dtype=torch.float64, device=dev, requires_grad=True)
    rhob = torch.tensor(w0[3 + 2 * rp.ndim], dtype=torch.float64, device=dev)
    loss = driver_loss(eng, amt, w0t, nsteps, line_vz2_loss(eng, rhob, i_line), every=every,
                       clock0=(rp.it0, rp.qt0))
    loss.backward()
    out = (float(loss.item()), w0t.grad.detach().cpu().numpy().reshape(k, -1), amt.grad.detach().cpu().numpy())
    eng.close()
    return out


def check(rp, w0, pos, amap, nsteps, i_line, every=2, seed=7, check_amap=True):
    print(f"[adjoint] {rp.ndim}D {rp.n} rkon={rp.rkon} terms={rp.terms:#x} steps={nsteps}", flush=True)
    loss, gw, ga = gpu_grads(rp, w0, pos, amap, nsteps, i_line, every)
    fd = ReplayFD(oracle, rp, w0, pos, amap, nsteps, i_line)
    try:
        assert abs(loss - fd.base) <= 1e-10 * abs(fd.base), (loss, fd.base)
        rng = np.random.default_rng(seed)
        k = P.npert(rp.ndim)
        inner = interior_mask(rp)
        # initial-state direction: noise scaled by each field's size
        scale = np.sqrt((w0[:k, inner] ** 2).mean(axis=1)) + 1e-30
        delta = rng.standard_normal((k, rp.ncells)) * scale[:, None] * inner[None, :]
        got = float((gw * delta).sum())
        want = fd.dir_w0(delta, EPS)
        print(f"  w0 direction: adjoint {got:.12e} fd {want:.12e}", flush=True)
        assert abs(got - want) <= TOL * abs(want), (got, want)
        if check_amap and amap is not None:
            da = rng.standard_normal(amap.size) * np.abs(amap).mean()
            got = float((ga * da).sum())
            want = fd.dir_amap(da, EPS)
            print(f"  amap direction: adjoint {got:.12e} fd {want:.12e}", flush=True)
            assert abs(got - want) <= TOL * abs(want), (got, want)
    finally:
        fd.close()
    return gw, ga


def strat_map_case(ni, ndim, rkon, terms=P.TERMS_ALL):
    rp, w, pos = strat_case(ni=ni, ndim=ndim, rkon=rkon)
    rp.terms = terms
    ncol = rp.ncells // rp.n[0]
    amap = np.random.default_rng(3).uniform(150.0, 250.0, ncol)
    return rp, w, pos, amap


@pytest.mark.parametrize("rkon", [1, 0])
def test_adjoint_strat2d(rkon):
    rp, w, pos, amap = strat_map_case((16, 12), 2, rkon)
    check(rp, w, pos, amap, nsteps=4, i_line=7)


def test_adjoint_strat3d_rk4():
    rp, w, pos, amap = strat_map_case((12, 8, 8), 3, 1)
    check(rp, w, pos, amap, nsteps=3, i_line=6)


@pytest.mark.parametrize("group", ["flux", "eterm", "grav", "hrho", "he", "hmom", "hb"])
def test_adjoint_term_groups_2d(group):
    t = {"flux": P.T_FLUX, "eterm": P.T_ETERM, "grav": P.T_GRAV, "hrho": P.T_HRHO, "he": P.T_HE,
         "hmom": P.T_HMOM, "hb": P.T_HB}[group]
    rp, w, pos, amap = strat_map_case((16, 12), 2, 1, terms=t | P.T_SOURCE)
    check(rp, w, pos, amap, nsteps=3, i_line=7)


@pytest.mark.parametrize("group", ["flux", "hmom", "hb"])
def test_adjoint_term_groups_3d(group):
    t = {"flux": P.T_FLUX | P.T_ETERM | P.T_GRAV, "hmom": P.T_HMOM | P.T_HRHO | P.T_HE, "hb": P.T_HB}[group]
    rp, w, pos, amap = strat_map_case((12, 8, 8), 3, 1, terms=t | P.T_SOURCE)
    check(rp, w, pos, amap, nsteps=2, i_line=6)


def test_adjoint_orszag_tang_periodic():
    rp, w, pos = ot_case(ni=16, rkon=1)
    rp.moddton = 0
    rp.dt = 2e-3
    check(rp, w, pos, None, nsteps=3, i_line=9, check_amap=False)


def test_adjoint_checkpoint_interval_invariance():
    """The gradient does not depend on the checkpoint interval."""
    rp, w, pos, amap = strat_map_case((16, 12), 2, 1)
    ref = gpu_grads(rp, w, pos, amap, 5, 7, 5)
    for every in (1, 2, 3):
        got = gpu_grads(rp, w, pos, amap, 5, 7, every)
        assert got[0] == ref[0]
        assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2]), every

"""The oracle features the adjoint checks rest on (CPU only): the driver
amplitude map, the restart, and the record/replay of the hyperdiffusion
coefficients that turns the oracle into the frozen-coefficient map whose
derivative the adjoint computes."""
from __future__ import annotations

import numpy as np
import pytest

from smaug_amd import params as P
from tests.adjoint_cases import ReplayFD, interior_mask, line_loss_np
from tests.cases import strat_case

oracle = pytest.importorskip("oracle.pyoracle")


def _map_of_scalar_driver(rp, pos):
    """amap(j) = A exp(-(x_j - x0)^2 / sxy^2), the scalar driver's horizontal factor."""
    n0 = rp.n[0]
    x = pos[1].reshape(-1, n0)[:, 0]
    return rp.drv[0] * np.exp(-(x - rp.drv[3]) ** 2 / rp.drv[6] ** 2)


def test_driver_map_reproduces_scalar_driver():
    rp, w, pos = strat_case(ni=(16, 12), ndim=2, rkon=1)
    a = oracle.Oracle(rp, w, pos)
    a.step(3)
    ref = a.state()
    wdpos, _ = a.grid()
    a.close()
    b = oracle.Oracle(rp, w, pos)
    b.set_driver_map(_map_of_scalar_driver(rp, wdpos))
    b.step(3)
    got = b.state()
    b.close()
    k = P.npert(2)
    d = np.abs(got[:k] - ref[:k]).max(axis=1) / np.abs(ref[:k]).max(axis=1)
    assert np.all(d < 1e-13), d
    # and a different map changes the momentum
    c = oracle.Oracle(rp, w, pos)
    c.set_driver_map(2.0 * _map_of_scalar_driver(rp, wdpos))
    c.step(1)
    assert not np.allclose(c.state()[1], ref[1])
    c.close()


def test_restart_and_replay_reproduce_the_run():
    rp, w, pos = strat_case(ni=(16, 12), ndim=2, rkon=1)
    o = oracle.Oracle(rp, w, pos)
    o.nu_mode(o.NU_RECORD)
    o.step(3)
    first = o.state()
    npairs = o.nu_mode(o.NU_REPLAY)
    # 3 steps x 4 stages x (2 rho + 2 e + 4 mom + 2 b) coefficient pairs in 2D
    assert npairs == 3 * 4 * 10
    o.restart(w, rp.it0, rp.qt0)
    o.step(3)
    assert np.array_equal(o.state(), first)
    o.close()


def test_replayed_coefficients_are_frozen():
    """With replay on, the coefficients do not follow the state: a state
    scaled by 1 + 1e-3 gives a different result than a recomputation."""
    rp, w, pos = strat_case(ni=(16, 12), ndim=2, rkon=1)
    k = P.npert(2)
    w2 = w.copy()
    w2[:k] *= 1.001
    o = oracle.Oracle(rp, w, pos)
    o.nu_mode(o.NU_RECORD)
    o.step(2)
    o.nu_mode(o.NU_REPLAY)
    o.restart(w2, rp.it0, rp.qt0)
    o.step(2)
    replayed = o.state()
    o.nu_mode(o.NU_OFF)
    o.restart(w2, rp.it0, rp.qt0)
    o.step(2)
    recomputed = o.state()
    o.close()
    assert not np.array_equal(replayed[:k], recomputed[:k])


def test_replay_fd_is_a_linear_map():
    """Extrapolated central differences of the frozen-coefficient map do not
    depend on the step (h and 2h agree) and are additive in the direction:
    the map is smooth and the differences resolve its derivative."""
    rp, w, pos = strat_case(ni=(16, 12), ndim=2, rkon=1)
    ncol = rp.ncells // rp.n[0]
    amap = np.random.default_rng(3).uniform(150.0, 250.0, ncol)
    fd = ReplayFD(oracle, rp, w, pos, amap, 3, 7)
    try:
        rng = np.random.default_rng(5)
        inner = interior_mask(rp)
        k = P.npert(2)
        scale = np.sqrt((w[:k, inner] ** 2).mean(axis=1))
        d1 = rng.standard_normal((k, rp.ncells)) * scale[:, None] * inner
        d2 = rng.standard_normal((k, rp.ncells)) * scale[:, None] * inner
        a = fd.dir_w0(d1, 1e-6)
        b = fd.dir_w0(d2, 1e-6)
        ab = fd.dir_w0(d1 + d2, 1e-6)
        assert abs(fd.dir_w0(d1, 2e-6) - a) <= 1e-7 * abs(a)
        assert abs(ab - (a + b)) <= 1e-7 * max(abs(a), abs(b))
        da = rng.standard_normal(ncol) * 200.0
        g = fd.dir_amap(da, 1e-6)
        assert abs(fd.dir_amap(da, 2e-6) - g) <= 1e-7 * abs(g)
        assert g != 0.0
    finally:
        fd.close()


def test_line_loss_reads_interior_line():
    rp, w, pos = strat_case(ni=(16, 12), ndim=2, rkon=1)
    n0, n1 = rp.n[0], rp.n[1]
    v = w.reshape(w.shape[0], n1, n0)
    i = 7
    want = sum((v[1, j, i] / (v[0, j, i] + v[7, j, i])) ** 2 for j in range(2, n1 - 2))
    assert line_loss_np(rp, w, i) == pytest.approx(want, rel=1e-14)

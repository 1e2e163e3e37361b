"""Shared parity cases: small configurations the oracle finishes in seconds."""
from __future__ import annotations

import numpy as np

from smaug_amd import configs
from smaug_amd.params import npert


def ot_case(ni=16, rkon=0):
    rp = configs.orszag_tang_params(ni=ni, rkon=rkon)
    w, pos = configs.orszag_tang(rp)
    return rp, w, pos


def strat_case(ni=(14, 10, 12), ndim=3, rkon=0, seed=20261202, driver=True):
    rp = configs.stratified_params(ni=ni[:ndim], ndim=ndim, rkon=rkon)
    if not driver:
        rp.driver = 0
    else:
        # driver footprint scaled to the small grid (a few cells wide, centred
        # near the bottom) and phase near its maximum, so the source term is
        # exercised at full strength rather than through exp(-35) tails
        dx = rp.dx
        rp.drv = [200.0, 300.0, rp.xmin[0] + 2.5 * dx[0], 2.0e6, 2.0e6, 3.0 * dx[0], 3.0 * dx[1], 0.0]
        rp.qt0 = 75.0
    w, pos = configs.stratified(rp, perturb=seed, amp=1e-3)
    return rp, w, pos


def rel_l2(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Per-variable relative L2 of a against reference b (rows = variables)."""
    num = np.sqrt(((a - b) ** 2).sum(axis=1))
    den = np.sqrt((b ** 2).sum(axis=1))
    den = np.where(den == 0, 1.0, den)
    return num / den


def compare(rp, w_hip, w_ref):
    k = npert(rp.ndim)
    return rel_l2(w_hip[:k], w_ref[:k])

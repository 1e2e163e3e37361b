"""Parity at the benchmark configurations' own sizes (GPU vs the oracle).

* Orszag-Tang vortex exactly as configured: 128 x 128 on periodic [0,1]^2,
  gamma 5/3, RK4, CFL 0.5 (PMODDTON 1), 200 steps, all variables;
* the stratified atmosphere at 128^3 (the 3D FP64 configuration's grid) for
  two RK4 steps -- the oracle needs ~15 s per step there, so the 1000-step
  run is checked on smaller grids (test_gpu_parity.py) and here only at
  full size.
Bar: relative L2 per evolved variable <= 1e-12, and identical step counts
and clocks.
"""
import time

import numpy as np
import pytest

from smaug_amd import configs
from smaug_amd.engine import Engine
from tests.cases import compare

oracle = pytest.importorskip("oracle.pyoracle")

pytestmark = pytest.mark.gpu

TOL = 1e-12


def test_orszag_tang_128_rk4_200_steps():
    rp = configs.orszag_tang_params(ni=128, rkon=1, courant=0.5)
    w, pos = configs.orszag_tang(rp)
    t0 = time.time()
    print("[config] OT 128^2 RK4 CFL 0.5, 200 steps", flush=True)
    e = Engine(rp, w, pos)
    e.step(200)
    got = e.state()
    cg = e.clock()
    e.close()
    o = oracle.Oracle(rp, w, pos)
    for k in range(4):
        o.step(50)
        print(f"  oracle {50 * (k + 1)} steps, {time.time() - t0:.0f} s", flush=True)
    ref = o.state()
    co = o.clock()
    o.close()
    err = compare(rp, got, ref)
    print("rel L2 per variable:", err, "clock", cg, co, flush=True)
    assert cg[0] == co[0] == 200
    assert cg[1] == pytest.approx(co[1], rel=1e-12) and cg[3] == pytest.approx(co[3], rel=1e-12)
    assert np.all(err <= TOL), err


def test_stratified_128cubed_two_rk4_steps():
    rp = configs.stratified_params(ni=(128, 128, 128), rkon=1)
    w, pos = configs.stratified(rp, perturb=20261202)
    print("[config] stratified 128^3 RK4, 2 steps", flush=True)
    e = Engine(rp, w, pos)
    e.step(2)
    got = e.state()
    e.close()
    t0 = time.time()
    o = oracle.Oracle(rp, w, pos)
    for k in range(2):
        o.step(1)
        print(f"  oracle step {k + 1}, {time.time() - t0:.0f} s", flush=True)
    ref = o.state()
    o.close()
    err = compare(rp, got, ref)
    print("rel L2 per variable:", err, flush=True)
    assert np.all(err <= TOL), err

"""Boundary types beyond periodic and contcd4: cont (3), fixed (5), symm (6)
and asymm (7), per field and direction (the boundtype[field][dir] table).

CPU: the oracle's ghost fill checked against the defining property of each
type on cells that are ghosts in one direction only (src/gradops.cu:
bc3_cont_dir 947-1060, bc3_fixed_dir 1234-1268, bc3_symm_dir 1325-1395,
bc3_asymm_dir 1397-1466):
  cont   ghost = nearest interior cell along the direction
  fixed  ghost = its value in the initial state
  symm   mirror (3D: 0<-4, 1<-3; 2D: 0<-3, 1<-2, likewise at the top)
  asymm  minus the symm value
and, on corner cells, the reference's pass order (direction 0, 1, 2, the
last pass in which a cell is a ghost decides).  GPU: the HIP engine against
the oracle with these tables over several RK4 and Euler steps (1e-12).
"""
from __future__ import annotations

import numpy as np
import pytest

from smaug_amd import params as P
from tests.cases import compare, strat_case

oracle = pytest.importorskip("oracle.pyoracle")

CONT, FIXED, SYMM, ASYMM = P.BC_CONT, P.BC_FIXED, P.BC_SYMM, P.BC_ASYMM


def mirror_src(t, x, n, nd):
    if t == CONT:
        return 2 if x < 2 else n - 3
    if nd == 2:
        return 3 - x if x < 2 else 2 * n - 5 - x
    return 4 - x if x < 2 else 2 * n - 6 - x


def bc_case(nd, rkon, types, field_types=None):
    ni = (16, 12) if nd == 2 else (12, 8, 8)
    rp, w, pos = strat_case(ni=ni, ndim=nd, rkon=rkon)
    rp.bcu = list(types) + [4] * (3 - nd)
    rp.bcl = list(rp.bcu)
    rp.bc_field = dict(field_types or {})
    # the initial state's ghost cells carry their own noise, so a fixed
    # boundary is visible against any image rule
    rng = np.random.default_rng(11)
    k = P.npert(nd)
    shape = tuple(rp.n[::-1])
    ghost = np.zeros(shape, dtype=bool)
    for d in range(nd):
        sl = [slice(None)] * nd
        ax = nd - 1 - d
        sl[ax] = np.r_[0, 1, rp.n[d] - 2, rp.n[d] - 1]
        ghost[tuple(sl)] = True
    g = ghost.ravel()
    for f in range(k):  # relative to the local perturbation, as small as it
        w[f, g] *= 1.0 + 1e-2 * rng.standard_normal(g.sum())
    return rp, w, pos


def types_of(rp, f):
    names = P.field_names(rp.ndim)
    return rp.bc_field.get(names[f], rp.bcu[: rp.ndim])


CASES = {
    "2d_symm_cont": (2, [SYMM, CONT], {"mom1": [ASYMM, CONT]}),
    "2d_fixed_asymm": (2, [FIXED, ASYMM], {"b2": [FIXED, SYMM]}),
    "3d_cont_symm_fixed": (3, [CONT, SYMM, FIXED], {"mom2": [CONT, ASYMM, FIXED]}),
    "3d_asymm_fixed_cont": (3, [ASYMM, FIXED, CONT], {"rho": [SYMM, FIXED, CONT], "b3": [CONT, CONT, ASYMM]}),
}


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_fill_properties(name):
    nd, types, ft = CASES[name]
    rp, w, pos = bc_case(nd, 1, types, ft)
    o = oracle.Oracle(rp, w, pos)
    got = o.state()
    o.close()
    n = rp.n
    shape = tuple(n[::-1])
    for f in range(P.npert(nd)):
        t = types_of(rp, f)
        a = got[f].reshape(shape)
        a0 = w[f].reshape(shape)
        it = np.ndindex(*shape)
        for idx in it:
            x = idx[::-1]  # (i, j[, k])
            gh = [d for d in range(nd) if x[d] < 2 or x[d] > n[d] - 3]
            if len(gh) != 1:
                continue
            d = gh[0]
            if t[d] == FIXED:
                assert a[idx] == a0[idx], (f, x)
                continue
            y = list(x)
            y[d] = mirror_src(t[d], x[d], n[d], nd)
            want = a[tuple(y[::-1])] * (-1.0 if t[d] == ASYMM else 1.0)
            assert a[idx] == want, (f, x, t)


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_corner_pass_order(name):
    """A cell that is a ghost in several directions takes the value of the
    last direction's pass (directions run 0, 1, 2)."""
    nd, types, ft = CASES[name]
    rp, w, pos = bc_case(nd, 1, types, ft)
    o = oracle.Oracle(rp, w, pos)
    got = o.state()
    o.close()
    n = rp.n
    shape = tuple(n[::-1])
    checked = 0
    for f in range(P.npert(nd)):
        t = types_of(rp, f)
        a = got[f].reshape(shape)
        a0 = w[f].reshape(shape)
        for idx in np.ndindex(*shape):
            x = list(idx[::-1])
            gh = [d for d in range(nd) if x[d] < 2 or x[d] > n[d] - 3]
            if len(gh) < 2:
                continue
            last = gh[-1]
            if t[last] == FIXED:
                assert a[idx] == a0[idx]
            else:
                y = list(x)
                y[last] = mirror_src(t[last], x[last], n[last], nd)
                assert a[idx] == a[tuple(y[::-1])] * (-1.0 if t[last] == ASYMM else 1.0)
            checked += 1
    assert checked > 0


@pytest.mark.gpu
@pytest.mark.parametrize("rkon", [1, 0])
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_boundaries_match_oracle(name, rkon):
    from smaug_amd.engine import Engine

    nd, types, ft = CASES[name]
    rp, w, pos = bc_case(nd, rkon, types, ft)
    print(f"[bc] {name} rkon={rkon}", flush=True)
    e = Engine(rp, w, pos)
    first = e.state()
    e.step(5)
    got = e.state()
    e.close()
    o = oracle.Oracle(rp, w, pos)
    ref0 = o.state()
    o.step(5)
    ref = o.state()
    o.close()
    assert np.all(compare(rp, first, ref0) <= 1e-15)
    err = compare(rp, got, ref)
    print("rel L2 per variable:", err)
    assert np.all(err <= 1e-12), err

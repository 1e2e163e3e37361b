"""TEST INFRASTRUCTURE: runs one check of the engine compiled for the host
(tests/hostemu/build.py) and prints its result as one JSON line.  Started
as a subprocess by tests/test_hostemu.py with SMAUG_HIP_LIB pointing at the
host build, so the product binding (smaug_amd.engine) drives the same C ABI
and kernels the GPU build compiles.

  python -m tests.hostemu.check fwd  <case>            forward vs oracle
  python -m tests.hostemu.check mixed <case>           mixed precision vs oracle
  python -m tests.hostemu.check adj  <case> <terms>    adjoint vs oracle FD
  python -m tests.hostemu.check autograd <rkon>        smaug_amd.adjoint path
  python -m tests.hostemu.check tl   <case> <terms>    tangent-linear run vs oracle FD
                                                       and vs the adjoint (dot products)

case: s2r (2D stratified RK4), s2e (2D Euler), s3r (3D RK4), s3e (3D Euler,
mixed only), ot (2D periodic).
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from smaug_amd import params as P  # noqa: E402
from smaug_amd.engine import Engine  # noqa: E402
from tests.adjoint_cases import ReplayFD, interior_mask, tl_directional  # noqa: E402
from tests.cases import compare, ot_case, strat_case  # noqa: E402
import oracle.pyoracle as oracle  # noqa: E402


def make_case(name: str, terms: int = P.TERMS_ALL):
    if name.startswith("bc:"):  # bc:<tests.test_boundaries case>:<rkon>
        from tests.test_boundaries import CASES, bc_case

        _, key, rkon = name.split(":")
        nd, types, ft = CASES[key]
        rp, w, pos = bc_case(nd, int(rkon), types, ft)
        amap = np.random.default_rng(3).uniform(150.0, 250.0, rp.ncells // rp.n[0])
        rp.terms = terms
        return rp, w, pos, amap, (7 if nd == 2 else 6), (3 if nd == 2 else 2)
    if name == "s2d":  # 2D stratified RK4 with the divB source terms and a divergent b1
        rp, w, pos = strat_case(ni=(16, 12), ndim=2, rkon=1)
        rp.divbon, rp.divbfix = 1, 1
        z = pos[0]
        w = w.copy()
        w[4] += 1e-3 * np.sin(2 * np.pi * (z - z.min()) / (z.max() - z.min())) * w[8]
        rp.terms = terms
        return rp, w, pos, None, 7, 3
    if name in ("tor", "spi", "spe"):  # the torsional (3D RK4) and spicule (2D RK4, Euler) drivers
        from tests.test_drivers import spicule_case, torsional_case

        if name == "tor":
            rp, w, pos = torsional_case(rkon=1, terms=terms)
            return rp, w, pos, None, 6, 2
        rp, w, pos = spicule_case(ndim=2, rkon=1 if name == "spi" else 0, terms=terms)
        return rp, w, pos, None, 7, 3
    if name == "s3m":  # 3D RK4 on a grid of several stage-kernel regions and dim-2 chunks
        rp, w, pos = strat_case(ni=(30, 17, 37), ndim=3, rkon=1)
        rp.terms = terms
        return rp, w, pos, None, 6, 1
    if name == "ot":
        rp, w, pos = ot_case(ni=16, rkon=1)
        rp.moddton = 0
        rp.dt = 2e-3
        amap, il, nsteps = None, 9, 3
    else:
        nd = 3 if name.startswith("s3") else 2
        rkon = 0 if name.endswith("e") else 1
        rp, w, pos = strat_case(ni=(16, 12) if nd == 2 else (12, 8, 8), ndim=nd, rkon=rkon)
        amap = np.random.default_rng(3).uniform(150.0, 250.0, rp.ncells // rp.n[0])
        il, nsteps = (7, 3) if nd == 2 else (6, 2)
    rp.terms = terms
    return rp, w, pos, amap, il, nsteps


def caller_buffer(n: int, fill: float = 0.0) -> np.ndarray:
    """n doubles of caller memory for the C ABI, allocated by the host
    build's own hipMalloc, so the buffer sits against a guard page like the
    engine's buffers (EMU_ALIGN picks the leading or trailing one) and an
    access past either end of it ends the check."""
    import ctypes

    from smaug_amd import engine as E

    lib = E.load()
    alloc = lib._Z9hipMallocPPvm
    alloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    alloc.restype = ctypes.c_int
    ptr = ctypes.c_void_p()
    assert alloc(ctypes.byref(ptr), 8 * n) == 0 and ptr.value
    a = np.ctypeslib.as_array((ctypes.c_double * n).from_address(ptr.value))
    a[:] = fill
    return a


def engine_adjoint(rp, w0, pos, amap, nsteps, i_line):
    """Taped steps and adjoint steps through the C ABI; the caller's buffers
    (tapes, end state, lambda, map gradient, the driver map) are guarded
    host-build allocations standing in for device memory."""
    nd = rp.ndim
    k = P.npert(nd)
    e = Engine(rp, w0, pos)
    if amap is not None:
        am = caller_buffer(amap.size)
        am[:] = amap
        e.set_driver_map(device_ptr=am.ctypes.data)
    F = k * rp.ncells
    T = e.tape_doubles()
    tapes = [caller_buffer(T, np.nan) for _ in range(nsteps)]
    for n in range(nsteps):
        e.step_taped(tapes[n].ctypes.data)
    wend = caller_buffer(F)
    e.save_pert(wend.ctypes.data)
    lam = caller_buffer(F)
    gmap = caller_buffer(rp.ncells // rp.n[0])
    shape = tuple(rp.n[:nd][::-1])
    rhob = w0[3 + 2 * nd].reshape(shape)
    inner = tuple(slice(2, rp.n[d] - 2) for d in range(nd - 1, 0, -1)) + (i_line,)
    for n in reversed(range(nsteps)):
        if n == nsteps - 1:
            wn = wend
        else:
            wn = caller_buffer(F)
            wn[:] = tapes[n + 1][:F]
        v = wn.reshape((k,) + shape)
        rt = v[0][inner] + rhob[inner]
        vz = v[1][inner] / rt
        L = lam.reshape(v.shape)
        L[1][inner] += 2 * vz / rt
        L[0][inner] += -2 * vz * vz / rt
        e.adjoint_step(tapes[n].ctypes.data, wn.ctypes.data, lam.ctypes.data, gmap.ctypes.data)
    e.close()
    return lam.reshape(k, -1), gmap


def check_fwd(case: str):
    rp, w, pos, amap, _, nsteps = make_case(case)
    e = Engine(rp, w, pos)
    if amap is not None:
        e.set_driver_map(amap)
    e.step(nsteps)
    got = e.state()
    e.close()
    o = oracle.Oracle(rp, w, pos)
    if amap is not None:
        o.set_driver_map(amap)
    o.step(nsteps)
    ref = o.state()
    o.close()
    return {"rel_l2": compare(rp, got, ref).tolist()}


def check_mixed(case: str):
    """Mixed precision (FP32 shadows and stencil arithmetic) against the
    FP64 oracle, with the step count of the GPU mixed-precision tests."""
    if case == "ot":
        rp, w, pos = ot_case(ni=32, rkon=1)
        nsteps = 30
    elif case == "s2r":
        rp, w, pos = strat_case(ni=(24, 16), ndim=2, rkon=1)
        nsteps = 20
    else:
        rp, w, pos = strat_case(ni=(14, 10, 12), rkon=0 if case == "s3e" else 1)
        nsteps = 10
    rp.precision = P.PREC_MIXED
    e = Engine(rp, w, pos)
    e.step(nsteps)
    got = e.state()
    e.close()
    o = oracle.Oracle(rp, w, pos)
    o.step(nsteps)
    ref = o.state()
    o.close()
    return {"rel_l2": compare(rp, got, ref).tolist(), "finite": bool(np.all(np.isfinite(got)))}


def check_adj(case: str, terms: int):
    rp, w, pos, amap, il, nsteps = make_case(case, terms)
    gw, ga = engine_adjoint(rp, w, pos, amap, nsteps, il)
    fd = ReplayFD(oracle, rp, w, pos, amap, nsteps, il)
    rng = np.random.default_rng(7)
    k = P.npert(rp.ndim)
    inner = interior_mask(rp)
    scale = np.sqrt((w[:k, inner] ** 2).mean(axis=1)) + 1e-30
    out = {"fields": []}
    for f in range(k):  # one direction per evolved field
        delta = np.zeros((k, rp.ncells))
        delta[f] = rng.standard_normal(rp.ncells) * scale[f] * inner
        got, want = float((gw * delta).sum()), fd.dir_w0(delta, 1e-6)
        out["fields"].append([got, want])
    if amap is not None:
        da = rng.standard_normal(amap.size) * np.abs(amap).mean()
        out["amap"] = [float((ga * da).sum()), fd.dir_amap(da, 1e-6)]
    fd.close()
    return out


def check_tl(case: str, terms: int):
    """Directional derivatives of the line loss along the tangent-linear run
    (smaug_tl_*) against the replayed-coefficient differences of the oracle,
    one direction per evolved field and one in the driver map, and the same
    directions through the adjoint (<adjoint gradient, direction> equals the
    tangent-linear derivative up to roundoff: the two are transposes)."""
    rp, w, pos, amap, il, nsteps = make_case(case, terms)
    gw, ga = engine_adjoint(rp, w, pos, amap, nsteps, il)
    fd = ReplayFD(oracle, rp, w, pos, amap, nsteps, il)
    e = Engine(rp, w, pos)
    if amap is not None:
        e.set_driver_map(amap)
    rhob = w[3 + 2 * rp.ndim]
    rng = np.random.default_rng(7)
    k = P.npert(rp.ndim)
    inner = interior_mask(rp)
    scale = np.sqrt((w[:k, inner] ** 2).mean(axis=1)) + 1e-30
    out = {"fields": [], "adjoint": []}
    loss = None
    s0 = e.state()
    for f in range(k):
        delta = np.zeros((k, rp.ncells))
        delta[f] = rng.standard_normal(rp.ncells) * scale[f] * inner
        loss, tl = tl_directional(e, rp, rhob, delta, None, nsteps, il)
        out["fields"].append([tl, fd.dir_w0(delta, 1e-6)])
        out["adjoint"].append([tl, float((gw * delta).sum())])
    out["loss"] = [loss, fd.base]
    if amap is not None:
        da = rng.standard_normal(amap.size) * np.abs(amap).mean()
        _, tl = tl_directional(e, rp, rhob, np.zeros((k, rp.ncells)), da, nsteps, il)
        out["amap"] = [tl, fd.dir_amap(da, 1e-6)]
        out["amap_adjoint"] = [tl, float((ga * da).sum())]
    # the driver amplitude drv[0] (no map: the scalar amplitude is the control)
    if amap is None and rp.driver != P.DRIVER_NONE:
        _, tl = tl_directional(e, rp, rhob, np.zeros((k, rp.ncells)), None, nsteps, il, damp=rp.drv[0])
        out["amplitude"] = [tl, fd.dir_amplitude(1e-6)]
    # a direction of the static background (every field, ghosts included),
    # relative to each cell's value (the stratified background spans orders
    # of magnitude)
    nb = w.shape[0] - k
    dbg = rng.standard_normal((nb, rp.ncells)) * np.abs(w[k:])
    _, tl = tl_directional(e, rp, rhob, np.zeros((k, rp.ncells)), None, nsteps, il, dbg=dbg)
    out["bg"] = [tl, fd.dir_bg(dbg, 1e-6)]
    # the background's controls: photospheric temperature, transition height
    if case in ("s2r", "s2e", "s3r"):
        from smaug_amd import configs

        out["controls"] = {}
        for name in configs.STATE_CONTROLS:
            kw = dict(perturb=20261202, amp=1e-3)
            c0 = {"t_photo": 6000.0, "z_tr": 2.0e6}[name]
            t = configs.stratified_tangent(rp, name, **kw)
            _, tl = tl_directional(e, rp, rhob, t[:k], None, nsteps, il, dbg=t[k:])

            def state_of(eps, name=name, c0=c0, kw=kw):
                return configs.stratified(rp, **kw, **{name: c0 * (1.0 + eps)})[0]

            # d/d(relative change) = c0 d/d(control)
            out["controls"][name] = [tl * c0, fd.dir_state(state_of, 1e-6)]
    # the run leaves the engine's own state alone
    out["state_untouched"] = bool(np.array_equal(e.state(), s0))
    e.close()
    fd.close()
    return out


def check_tl_refusals():
    """The tangent-linear entry points refuse what they do not cover; from a
    mixed-precision engine the run's values and tangents stay within 1e-5 of
    the FP64 run's (and differ from them: the FP32 pairs are in use)."""
    from smaug_amd.engine import SmaugError

    out = {}

    def refused(name, fn):
        try:
            fn()
            out[name] = "accepted"
        except SmaugError as err:
            out[name] = str(err)

    rp, w, pos, _, _, _ = make_case("s2d")  # divB source terms on
    e = Engine(rp, w, pos)
    k = P.npert(rp.ndim)
    refused("divb", lambda: e.tl_begin(np.zeros((k, rp.ncells))))
    e.close()
    rp, w, pos = strat_case(ni=(16, 12), ndim=2, rkon=1)
    e = Engine(rp, w, pos)
    refused("step_before_begin", lambda: e.tl_step(1))
    refused("amplitude_before_begin", lambda: e.tl_driver_amplitude(1.0))
    e.tl_begin(np.zeros((k, rp.ncells)))
    refused("amplitude_not_finite", lambda: e.tl_driver_amplitude(float("nan")))
    e.set_driver_map(np.full(rp.ncells // rp.n[0], 200.0))
    e.tl_begin(np.zeros((k, rp.ncells)))
    refused("amplitude_with_map", lambda: e.tl_driver_amplitude(1.0))
    e.tl_end()
    e.set_driver_map(None)
    ro, wo, po = ot_case(ni=16, rkon=1)  # no driver
    eo = Engine(ro, wo, po)
    eo.tl_begin(np.zeros((P.npert(2), ro.ncells)))
    refused("amplitude_without_driver", lambda: eo.tl_driver_amplitude(1.0))
    eo.tl_end()
    eo.close()
    refused("map_direction_without_map",
            lambda: e.tl_begin(np.zeros((k, rp.ncells)), damap=np.ones(rp.ncells // rp.n[0])))
    inner = interior_mask(rp)
    rng = np.random.default_rng(9)
    scale = np.sqrt((w[:k, inner] ** 2).mean(axis=1))
    d = rng.standard_normal((k, rp.ncells)) * scale[:, None] * inner[None, :]
    runs = []
    for prec in (P.PREC_FP64, P.PREC_MIXED):
        rp.precision = prec
        e = Engine(rp, w, pos)
        e.tl_begin(d)
        e.tl_step(10)
        runs.append(e.tl_get())
        e.tl_end()
        e.close()
    rel = lambda a, b: (np.sqrt(((a - b) ** 2).sum(axis=1)) / np.sqrt((b ** 2).sum(axis=1))).tolist()
    out["mixed_values_rel_l2"] = rel(runs[1][0], runs[0][0])
    out["mixed_tangents_rel_l2"] = rel(runs[1][1], runs[0][1])
    return out


def check_tangent_torch():
    """smaug_amd.tangent.state_loss under torch's forward-mode AD on CPU
    tensors (the host build's device memory is host memory): value and
    tangents in a direction of the state and background and in the driver
    amplitude, against the oracle's loss and replayed differences."""
    import torch
    import torch.autograd.forward_ad as fwAD

    from smaug_amd.tangent import line_vz2_loss_bg, state_loss

    rp, w, pos = strat_case(ni=(16, 12), ndim=2, rkon=1)
    k, il, nsteps = P.npert(2), 7, 3
    rng = np.random.default_rng(5)
    inner = interior_mask(rp)
    scale = np.sqrt((w[:k, inner] ** 2).mean(axis=1))
    d = rng.standard_normal((k, rp.ncells)) * scale[:, None] * inner[None, :]
    dbg = rng.standard_normal(w[k:].shape) * np.abs(w[k:])
    w0, bg = torch.tensor(w[:k].ravel()), torch.tensor(w[k:].ravel())
    a0 = torch.tensor(rp.drv[0], dtype=torch.float64)
    e = Engine(rp, w, pos)
    fd = ReplayFD(oracle, rp, w, pos, None, nsteps, il)
    fn = line_vz2_loss_bg(e, il)
    with fwAD.dual_level():
        val, tan = fwAD.unpack_dual(state_loss(e, fwAD.make_dual(w0, torch.tensor(d.ravel())),
                                               fwAD.make_dual(bg, torch.tensor(dbg.ravel())), nsteps, fn,
                                               clock0=(rp.it0, rp.qt0), amplitude=fwAD.make_dual(a0, a0.clone())))
    out = {"loss": [float(val), fd.base],
           "all": [float(tan), fd.dir_w0(d, 1e-6) + fd.dir_bg(dbg, 1e-6) + fd.dir_amplitude(1e-6)]}
    try:
        state_loss(e, w0, bg, nsteps, fn, amplitude=1.5 * a0)
        out["wrong_amplitude"] = "accepted"
    except ValueError as err:
        out["wrong_amplitude"] = str(err)
    e.close()
    fd.close()
    return out


def check_autograd(rkon: int):
    """smaug_amd.adjoint.driver_loss on CPU tensors (the host build's device
    memory is host memory); torch's stream calls are replaced by no-ops."""
    import contextlib

    import torch

    class _Stream:
        cuda_stream = 0x1

        def wait_stream(self, other):
            pass

        def synchronize(self):
            pass

    st = _Stream()
    torch.cuda.current_stream = lambda *a, **k: st
    torch.cuda.Stream = lambda *a, **k: st
    torch.cuda.stream = lambda s: contextlib.nullcontext()
    torch.cuda.synchronize = lambda *a, **k: None
    torch.Tensor.record_stream = lambda self, s: None
    torch.Tensor.is_cuda = property(lambda self: True)
    from smaug_amd.adjoint import driver_loss, line_vz2_loss

    rp, w0, pos = strat_case(ni=(16, 12), ndim=2, rkon=rkon)
    amap = np.random.default_rng(3).uniform(150.0, 250.0, rp.ncells // rp.n[0])
    k = P.npert(2)
    nsteps, il = 5, 7
    res = {}
    for every in (5, 2, 1):
        eng = Engine(rp, w0, pos)
        w0t = torch.tensor(np.ascontiguousarray(w0[:k]).ravel(), dtype=torch.float64, requires_grad=True)
        amt = torch.tensor(amap, dtype=torch.float64, requires_grad=True)
        rhob = torch.tensor(w0[3 + 4], dtype=torch.float64)
        loss = driver_loss(eng, amt, w0t, nsteps, line_vz2_loss(eng, rhob, il), every=every,
                           clock0=(rp.it0, rp.qt0))
        loss.backward()
        res[every] = (float(loss.item()), w0t.grad.numpy().copy(), amt.grad.numpy().copy())
        eng.close()
    fd = ReplayFD(oracle, rp, w0, pos, amap, nsteps, il)
    rng = np.random.default_rng(7)
    inner = interior_mask(rp)
    scale = np.sqrt((w0[:k, inner] ** 2).mean(axis=1)) + 1e-30
    delta = rng.standard_normal((k, rp.ncells)) * scale[:, None] * inner[None, :]
    da = rng.standard_normal(amap.size) * amap.mean()
    gw, ga = res[2][1].reshape(k, -1), res[2][2]
    out = {
        "loss": res[2][0], "oracle_loss": fd.base,
        "w0": [float((gw * delta).sum()), fd.dir_w0(delta, 1e-6)],
        "amap": [float((ga * da).sum()), fd.dir_amap(da, 1e-6)],
        "every_identical": all(np.array_equal(res[e][1], res[5][1]) and np.array_equal(res[e][2], res[5][2])
                               for e in (2, 1)),
    }
    fd.close()
    return out


def check_blowup(kind: str):
    """A NaN in one cell's energy, or a negative total density, is reported
    as SMAUG_E_NUMERIC by the next synchronising call; a clean state loaded
    afterwards steps without the error."""
    from smaug_amd.engine import SmaugError

    rp, w, pos = strat_case(ni=(16, 12), ndim=2, rkon=1)
    nd = rp.ndim
    bad = w.copy()
    c = (6 * rp.n[0]) + 7  # an interior cell
    if kind == "nan":
        bad[1 + nd, c] = np.nan
    else:  # rho + rho_b < 0
        bad[0, c] = -2.0 * w[3 + 2 * nd, c]
    e = Engine(rp, bad, pos)
    e.step(1)
    out = {"kind": kind}
    try:
        e.sync()
        out["raised"] = None
    except SmaugError as err:
        out["raised"] = str(err)
    e.set_state(w)
    e.step(1)
    try:
        e.sync()
        out["after_reset"] = "ok"
    except SmaugError as err:
        out["after_reset"] = str(err)
    e.close()
    return out


def check_noalloc():
    """smaug_cfl_noalloc against smaug_cfl on two states in a row, the
    second on a smaller grid, the states in guarded caller buffers."""
    from smaug_amd.engine import cfl_noalloc

    out = []
    for ni, nd, moddton in (((14, 10, 8), 3, 1), ((12, 8), 2, 0)):
        rp, w, pos = strat_case(ni=ni, ndim=nd, rkon=1)
        rp.moddton = moddton
        rp.courant = 0.3
        if moddton:
            rp.dt = 10.0  # above the CFL step, so the rule picks courant / courantmax
        e = Engine(rp, w, pos)
        cm, dt = e.cfl()
        state = e.state()
        e.close()
        wd = caller_buffer(state.size)
        wd[:] = state.ravel()
        res = caller_buffer(nd + 2, np.nan)
        cfl_noalloc(rp, wd.ctypes.data, res.ctypes.data)
        out.append({"ndim": nd, "cmax": list(map(float, cm)), "dt": float(dt),
                    "got": list(map(float, res))})
    return out


def check_ranges():
    """Caller device buffers shorter than the call needs are refused with
    SMAUG_E_ARG before anything is queued (the short buffers are guarded
    host-build allocations, so a kernel or copy that ran anyway would stop
    the check at the guard page)."""
    from smaug_amd.engine import SmaugError, cfl_noalloc

    rp, w, pos = strat_case(ni=(16, 12), ndim=2, rkon=1)
    k = P.npert(rp.ndim)
    F = k * rp.ncells
    e = Engine(rp, w, pos)
    out = {}

    def refused(name, fn):
        try:
            fn()
            out[name] = "accepted"
        except SmaugError as err:
            out[name] = str(err)

    short = caller_buffer(F - 1)
    refused("save_pert", lambda: e.save_pert(short.ctypes.data))
    refused("load_pert", lambda: e.load_pert(short.ctypes.data))
    tape = caller_buffer(e.tape_doubles() - 1)
    refused("step_taped", lambda: e.step_taped(tape.ctypes.data))
    good_tape = caller_buffer(e.tape_doubles())
    e.step_taped(good_tape.ctypes.data)
    wn = caller_buffer(F)
    e.save_pert(wn.ctypes.data)
    refused("adjoint_step", lambda: e.adjoint_step(good_tape.ctypes.data, wn.ctypes.data, short.ctypes.data))
    amap = caller_buffer(rp.ncells // rp.n[0] - 1)
    refused("set_driver_map", lambda: e.set_driver_map(device_ptr=amap.ctypes.data))
    state = caller_buffer(P.nvar(rp.ndim) * rp.ncells)
    state[:] = e.state().ravel()
    res = caller_buffer(rp.ndim + 1)
    refused("cfl_noalloc", lambda: cfl_noalloc(rp, state.ctypes.data, res.ctypes.data))
    e.close()
    return out


def main(argv):
    what = argv[0]
    if what == "ranges":
        r = check_ranges()
    elif what == "noalloc":
        r = check_noalloc()
    elif what == "blowup":
        r = check_blowup(argv[1])
    elif what == "fwd":
        r = check_fwd(argv[1])
    elif what == "mixed":
        r = check_mixed(argv[1])
    elif what == "adj":
        r = check_adj(argv[1], int(argv[2], 0) if len(argv) > 2 else P.TERMS_ALL)
    elif what == "tl_refusals":
        r = check_tl_refusals()
    elif what == "tl_torch":
        r = check_tangent_torch()
    elif what == "tl":
        r = check_tl(argv[1], int(argv[2], 0) if len(argv) > 2 else P.TERMS_ALL)
    else:
        r = check_autograd(int(argv[1]))
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])

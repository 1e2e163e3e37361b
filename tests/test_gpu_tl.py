"""Tangent-linear run (smaug_tl_*) on the MI355X: the stage kernel
instantiated on (value, tangent) pairs, with the hyperdiffusion coefficients,
the wave-speed maxima and dt held at their values along the trajectory.

Checked against the oracle with its hyperdiffusion coefficients replayed
from the unperturbed run (the derivative of the step with those
coefficients frozen), by Richardson-extrapolated central differences at
h = 1e-6 relative: the directional derivative of the summed line loss
sum_n sum_line (m_0 / rho_t)^2 agrees to 1e-6 relative, per evolved field.
The values of the run reproduce the oracle's loss (1e-12) and the FP64
forward step's state (1e-13 per variable), and the run leaves the engine's
own state untouched.  The driver amplitude map stays off here: it belongs
to the differentiable entry points gated off on the GPU (SMAUG_ADJOINT), so
its direction is checked in the host build (tests/test_hostemu.py).  The
driver's scalar amplitude drv[0] is a direction here
(smaug_tl_driver_amplitude), against the oracle rerun with its amplitude
moved.
"""
from __future__ import annotations

import numpy as np
import pytest

from smaug_amd import params as P
from tests.adjoint_cases import ReplayFD, interior_mask, tl_directional
from tests.cases import compare, ot_case, strat_case

oracle = pytest.importorskip("oracle.pyoracle")
torch = pytest.importorskip("torch")  # before the engine library loads: one HIP runtime

pytestmark = pytest.mark.gpu

TOL = 1e-6
EPS = 1e-6


def tl_check(rp, w, pos, nsteps, i_line, per_field=True, seed=7):
    from smaug_amd.engine import Engine

    print(f"[tl] {rp.ndim}D {rp.n} rkon={rp.rkon} terms={rp.terms:#x} driver={rp.driver} steps={nsteps}", flush=True)
    e = Engine(rp, w, pos)
    s0 = e.state()
    fd = ReplayFD(oracle, rp, w, pos, None, nsteps, i_line)
    try:
        k = P.npert(rp.ndim)
        rhob = w[3 + 2 * rp.ndim]
        inner = interior_mask(rp)
        rng = np.random.default_rng(seed)
        scale = np.sqrt((w[:k, inner] ** 2).mean(axis=1)) + 1e-30
        dirs = []
        if per_field:
            for f in range(k):
                d = np.zeros((k, rp.ncells))
                d[f] = rng.standard_normal(rp.ncells) * scale[f] * inner
                dirs.append(d)
        else:
            dirs.append(rng.standard_normal((k, rp.ncells)) * scale[:, None] * inner[None, :])
        for i, d in enumerate(dirs):
            loss, got = tl_directional(e, rp, rhob, d, None, nsteps, i_line)
            want = fd.dir_w0(d, EPS)
            print(f"  w0 direction {i}: tangent-linear {got:.12e} fd {want:.12e}", flush=True)
            assert abs(loss - fd.base) <= 1e-12 * abs(fd.base), (loss, fd.base)
            if want == 0.0:
                assert abs(got) <= 1e-12, (i, got)
            else:
                assert abs(got - want) <= TOL * abs(want), (i, got, want)
        # a direction of the background fields, relative to each cell's value
        dbg = rng.standard_normal(w[k:].shape) * np.abs(w[k:])
        if np.any(dbg):
            _, got = tl_directional(e, rp, rhob, np.zeros((k, rp.ncells)), None, nsteps, i_line, dbg=dbg)
            want = fd.dir_bg(dbg, EPS)
            print(f"  background direction: tangent-linear {got:.12e} fd {want:.12e}", flush=True)
            assert abs(got - want) <= TOL * abs(want), (got, want)
        # the run's values against the FP64 forward step of the same engine
        e.tl_begin(np.zeros((k, rp.ncells)))
        e.tl_step(nsteps)
        wv, dw = e.tl_get()
        e.tl_end()
        assert np.array_equal(e.state(), s0)
        assert not np.any(dw)
        e.step(nsteps)
        ref = e.state()
        full = ref.copy()
        full[:k] = wv
        assert max(compare(rp, full, ref)) <= 1e-13
    finally:
        fd.close()
        e.close()


def strat_tl_case(ni, ndim, rkon, terms=P.TERMS_ALL):
    rp, w, pos = strat_case(ni=ni, ndim=ndim, rkon=rkon)
    rp.terms = terms
    return rp, w, pos


@pytest.mark.parametrize("rkon", [1, 0])
def test_tl_strat2d(rkon):
    rp, w, pos = strat_tl_case((16, 12), 2, rkon)
    tl_check(rp, w, pos, nsteps=4, i_line=7)


def test_tl_strat3d_rk4():
    rp, w, pos = strat_tl_case((12, 8, 8), 3, 1)
    tl_check(rp, w, pos, nsteps=3, i_line=6)


@pytest.mark.parametrize("group", ["flux", "eterm", "grav", "hrho", "he", "hmom", "hb"])
def test_tl_term_groups_2d(group):
    t = {"flux": P.T_FLUX, "eterm": P.T_ETERM, "grav": P.T_GRAV, "hrho": P.T_HRHO, "he": P.T_HE,
         "hmom": P.T_HMOM, "hb": P.T_HB}[group]
    rp, w, pos = strat_tl_case((16, 12), 2, 1, terms=t | P.T_SOURCE)
    tl_check(rp, w, pos, nsteps=3, i_line=7)


@pytest.mark.parametrize("group", ["flux", "hmom", "hb"])
def test_tl_term_groups_3d(group):
    t = {"flux": P.T_FLUX | P.T_ETERM | P.T_GRAV, "hmom": P.T_HMOM | P.T_HRHO | P.T_HE, "hb": P.T_HB}[group]
    rp, w, pos = strat_tl_case((12, 8, 8), 3, 1, terms=t | P.T_SOURCE)
    tl_check(rp, w, pos, nsteps=2, i_line=6)


def test_tl_orszag_tang_periodic():
    """Periodic boundaries with the flux and background terms.  With
    hyperdiffusion on, the Orszag-Tang derivative with frozen coefficients is
    itself set by roundoff: the smooth initial data leave third differences
    at roundoff level, so nu along a trajectory started one ulp away differs
    at O(1) on some interfaces and the oracle's own replayed derivative moves
    by up to 4 % (16^2, 3 steps; 8e-13 on the stratified case).  That case is
    checked in the host build, whose arithmetic reproduces the oracle's nu."""
    rp, w, pos = ot_case(ni=16, rkon=1)
    rp.moddton = 0
    rp.dt = 2e-3
    rp.terms = P.TERMS_ALL & ~(P.T_HRHO | P.T_HE | P.T_HMOM | P.T_HB)
    tl_check(rp, w, pos, nsteps=3, i_line=9)


def test_tl_inner_tiles_3d():
    """A 3D grid with tiles away from every face (the stage kernel's inner
    body) as well as edge tiles: one direction over all fields."""
    rp, w, pos = strat_tl_case((30, 17, 37), 3, 1)
    tl_check(rp, w, pos, nsteps=2, i_line=9, per_field=False)


@pytest.mark.parametrize("case", ["tor", "spi"])
def test_tl_reference_drivers(case):
    from tests.test_drivers import spicule_case, torsional_case

    if case == "tor":
        rp, w, pos = torsional_case(rkon=1, terms=P.TERMS_ALL)
        tl_check(rp, w, pos, nsteps=2, i_line=6)
    else:
        rp, w, pos = spicule_case(ndim=2, rkon=1, terms=P.TERMS_ALL)
        tl_check(rp, w, pos, nsteps=3, i_line=7)


@pytest.mark.parametrize("key,rkon", [("2d_fixed_asymm", 0), ("3d_cont_symm_fixed", 1)])
def test_tl_boundary_types(key, rkon):
    from tests.test_boundaries import CASES, bc_case

    nd, types, ft = CASES[key]
    rp, w, pos = bc_case(nd, rkon, types, ft)
    tl_check(rp, w, pos, nsteps=3 if nd == 2 else 2, i_line=7 if nd == 2 else 6)


def test_tl_c4_slice_512x256():
    """The 2D adjoint configuration's grid, 512 (z) x 256 (x), 5 RK4 steps:
    one direction over all fields against the replayed differences."""
    from smaug_amd import configs

    rp = configs.stratified_params(ni=(512, 256), ndim=2, rkon=1)
    w, pos = configs.stratified(rp, perturb=20261202)
    tl_check(rp, w, pos, nsteps=5, i_line=340, per_field=False)


def test_tl_c5_grid_192cubed_properties():
    """192^3 (the 3D adjoint configuration's grid), 2 RK4 steps, where the
    oracle is too slow for differences: the tangents are linear in the
    direction (t(d1 + 2 d2) = t(d1) + 2 t(d2) to roundoff), scale exactly
    (bitwise) with a power of two, vanish for a zero direction, and the
    values equal the FP64 step's state."""
    from smaug_amd import configs
    from smaug_amd.engine import Engine

    rp = configs.stratified_params(ni=(192, 192, 192), rkon=1)
    w, pos = configs.stratified(rp, perturb=20261202)
    k = P.npert(3)
    inner = interior_mask(rp)
    rng = np.random.default_rng(11)
    scale = np.sqrt((w[:k, inner] ** 2).mean(axis=1)) + 1e-30
    d1 = rng.standard_normal((k, rp.ncells)) * scale[:, None] * inner[None, :]
    d2 = rng.standard_normal((k, rp.ncells)) * scale[:, None] * inner[None, :]
    e = Engine(rp, w, pos)
    try:
        def tangent(d, nsteps=2):
            e.tl_begin(d)
            e.tl_step(nsteps)
            wv, dw = e.tl_get()
            e.tl_end()
            return wv, dw

        print("[tl] 192^3 RK4: four tangent-linear runs of 2 steps", flush=True)
        wv, t1 = tangent(d1)
        _, t2 = tangent(d2)
        _, t3 = tangent(d1 + 2.0 * d2)
        lin = np.sqrt(((t3 - (t1 + 2.0 * t2)) ** 2).sum(axis=1) / (t3 ** 2).sum(axis=1))
        print("  linearity, rel L2 per variable:", lin, flush=True)
        assert np.all(lin <= 1e-12), lin
        _, t4 = tangent(4.0 * d1)
        assert np.array_equal(t4, 4.0 * t1)
        _, t0 = tangent(np.zeros_like(d1), nsteps=1)
        assert not np.any(t0)
        e.step(2)
        ref = e.state()
        full = ref.copy()
        full[:k] = wv
        err = compare(rp, full, ref)
        print("  values vs the FP64 step, rel L2 per variable:", err, flush=True)
        assert max(err) <= 1e-13, err
    finally:
        e.close()


@pytest.mark.parametrize("ni,nsteps,i_line", [((16, 12), 4, 7), ((12, 8, 8), 3, 6), ((512, 256), 5, 340)])
def test_tl_background_controls(ni, nsteps, i_line):
    """The derivative of the summed line loss with respect to the
    background's controls -- the photospheric temperature and the transition
    height of T(z), which set the hydrostatic density, pressure and energy
    and, through them, the seeded perturbation -- from one tangent-linear
    run each, in the direction of configs.stratified_tangent, against the
    oracle rerun from the state rebuilt with the control moved by +-1e-6 and
    +-2e-6 relative (coefficients replayed).  (512, 256) is the 2D adjoint
    configuration's grid."""
    from smaug_amd import configs
    from smaug_amd.engine import Engine

    nd = len(ni)
    if ni == (512, 256):
        rp = configs.stratified_params(ni=ni, ndim=2, rkon=1)
        kw = dict(perturb=20261202)
    else:
        rp, _, _ = strat_case(ni=ni, ndim=nd, rkon=1)
        kw = dict(perturb=20261202, amp=1e-3)
    w, pos = configs.stratified(rp, **kw)
    k = P.npert(nd)
    print(f"[tl] background controls {nd}D {rp.n} steps={nsteps}", flush=True)
    e = Engine(rp, w, pos)
    fd = ReplayFD(oracle, rp, w, pos, None, nsteps, i_line)
    try:
        for name, c0 in (("t_photo", 6000.0), ("z_tr", 2.0e6)):
            t = configs.stratified_tangent(rp, name, **kw)
            _, got = tl_directional(e, rp, w[3 + 2 * nd], t[:k], None, nsteps, i_line, dbg=t[k:])
            got *= c0  # per relative change of the control

            def state_of(eps, name=name, c0=c0):
                return configs.stratified(rp, **kw, **{name: c0 * (1.0 + eps)})[0]

            want = fd.dir_state(state_of, EPS)
            print(f"  {name}: tangent-linear {got:.12e} fd {want:.12e}", flush=True)
            assert abs(got - want) <= TOL * abs(want), (name, got, want)
    finally:
        fd.close()
        e.close()


def test_tl_torch_forward_mode():
    """smaug_amd.tangent.state_loss under torch's forward-mode AD (dual
    tensors on the GPU, the loss in torch): the value equals the oracle's
    loss and the tangent its replayed differences, in a direction of the
    initial state and of the background together, of the state alone and
    of the driver amplitude."""
    import torch.autograd.forward_ad as fwAD

    from smaug_amd.engine import Engine
    from smaug_amd.tangent import line_vz2_loss_bg, state_loss

    rp, w, pos = strat_tl_case((16, 12), 2, 1)
    k, il, nsteps = P.npert(2), 7, 4
    rng = np.random.default_rng(5)
    inner = interior_mask(rp)
    scale = np.sqrt((w[:k, inner] ** 2).mean(axis=1))
    d = rng.standard_normal((k, rp.ncells)) * scale[:, None] * inner[None, :]
    dbg = rng.standard_normal(w[k:].shape) * np.abs(w[k:])
    dev = torch.device("cuda:0")
    w0 = torch.tensor(w[:k].ravel(), device=dev)
    bg = torch.tensor(w[k:].ravel(), device=dev)
    e = Engine(rp, w, pos)
    e.step(3)  # state_loss starts from w0 whatever the engine holds
    fd = ReplayFD(oracle, rp, w, pos, None, nsteps, il)
    try:
        fn = line_vz2_loss_bg(e, il)
        with fwAD.dual_level():
            a = fwAD.make_dual(w0, torch.tensor(d.ravel(), device=dev))
            b = fwAD.make_dual(bg, torch.tensor(dbg.ravel(), device=dev))
            val, tan = fwAD.unpack_dual(state_loss(e, a, b, nsteps, fn, clock0=(rp.it0, rp.qt0)))
        want = fd.dir_w0(d, EPS) + fd.dir_bg(dbg, EPS)
        print(f"  state + background: forward mode {float(tan):.12e} fd {want:.12e}", flush=True)
        assert abs(float(val) - fd.base) <= 1e-12 * abs(fd.base), (float(val), fd.base)
        assert abs(float(tan) - want) <= TOL * abs(want), (float(tan), want)
        with fwAD.dual_level():
            a = fwAD.make_dual(w0, torch.tensor(d.ravel(), device=dev))
            tan = fwAD.unpack_dual(state_loss(e, a, bg, nsteps, fn, clock0=(rp.it0, rp.qt0))).tangent
        want = fd.dir_w0(d, EPS)
        assert abs(float(tan) - want) <= TOL * abs(want), (float(tan), want)
        # the driver amplitude as a dual input: tangent A gives A dL/dA
        a0 = torch.tensor(rp.drv[0], dtype=torch.float64, device=dev)
        with fwAD.dual_level():
            amp = fwAD.make_dual(a0, a0.clone())
            tan = fwAD.unpack_dual(state_loss(e, w0, bg, nsteps, fn, clock0=(rp.it0, rp.qt0),
                                              amplitude=amp)).tangent
        want = fd.dir_amplitude(EPS)
        print(f"  amplitude: forward mode {float(tan):.12e} fd {want:.12e}", flush=True)
        assert abs(float(tan) - want) <= TOL * abs(want), (float(tan), want)
        with pytest.raises(ValueError):  # the engine's amplitude is fixed at creation
            state_loss(e, w0, bg, nsteps, fn, amplitude=1.5 * a0)
    finally:
        fd.close()
        e.close()


@pytest.mark.parametrize("case", ["strat2d", "strat3d", "tor", "spi"])
def test_tl_driver_amplitude(case):
    """The derivative in the driver amplitude drv[0] (the vertical driver on
    2D and 3D stratified grids, the torsional and the spicule drivers), from
    one tangent-linear run with smaug_tl_driver_amplitude, against the
    oracle rerun with its amplitude moved by +-1e-6 and +-2e-6 relative
    (coefficients replayed)."""
    from smaug_amd.engine import Engine
    from tests.test_drivers import spicule_case, torsional_case

    if case == "strat2d":
        (rp, w, pos), nsteps, il = strat_tl_case((16, 12), 2, 1), 4, 7
    elif case == "strat3d":
        (rp, w, pos), nsteps, il = strat_tl_case((12, 8, 8), 3, 1), 3, 6
    elif case == "tor":
        (rp, w, pos), nsteps, il = torsional_case(rkon=1, terms=P.TERMS_ALL), 2, 6
    else:
        (rp, w, pos), nsteps, il = spicule_case(ndim=2, rkon=1, terms=P.TERMS_ALL), 3, 7
    k = P.npert(rp.ndim)
    print(f"[tl] driver amplitude {case} {rp.n} steps={nsteps}", flush=True)
    e = Engine(rp, w, pos)
    fd = ReplayFD(oracle, rp, w, pos, None, nsteps, il)
    try:
        loss, got = tl_directional(e, rp, w[3 + 2 * rp.ndim], np.zeros((k, rp.ncells)), None, nsteps, il,
                                   damp=rp.drv[0])
        want = fd.dir_amplitude(EPS)
        print(f"  A dL/dA: tangent-linear {got:.12e} fd {want:.12e}", flush=True)
        assert abs(loss - fd.base) <= 1e-12 * abs(fd.base), (loss, fd.base)
        assert abs(got - want) <= TOL * abs(want), (got, want)
    finally:
        fd.close()
        e.close()


@pytest.mark.parametrize("control", ["t_photo", "z_tr", "amplitude"])
def test_tl_slice_horizon_controls(control):
    """The 2D slice configuration over its loss's own horizon: 64 (z) x 32
    (x), driven from rest, 1500 RK4 steps, L = sum over the steps of v_z^2 on
    the line z ~ 4 Mm; dL/d(control) from one tangent-linear run against the
    oracle's replayed differences over the same 1500 steps, for the
    background's controls and the driver amplitude.  Moving the trajectory
    by one ulp moves the background controls' derivatives by 1.3e-9 (T_photo)
    and 2.9e-8 (z_tr) relative, so the 1e-6 bar holds above the roundoff
    floor."""
    from smaug_amd import configs
    from smaug_amd.engine import Engine

    rp = configs.stratified_params(ni=(64, 32), ndim=2, rkon=1)
    w, pos = configs.stratified(rp)
    k, nsteps, il = P.npert(2), 1500, 44
    print(f"[tl] slice 64x32, {nsteps} steps, control {control}", flush=True)
    e = Engine(rp, w, pos)
    fd = ReplayFD(oracle, rp, w, pos, None, nsteps, il)
    try:
        if control == "amplitude":  # A dL/dA
            loss, got = tl_directional(e, rp, w[3 + 2 * rp.ndim], np.zeros((k, rp.ncells)), None, nsteps, il,
                                       damp=rp.drv[0])
            want = fd.dir_amplitude(EPS)
        else:  # c0 dL/dc0
            c0 = {"t_photo": 6000.0, "z_tr": 2.0e6}[control]
            t = configs.stratified_tangent(rp, control)
            loss, got = tl_directional(e, rp, w[3 + 2 * rp.ndim], t[:k], None, nsteps, il, dbg=t[k:])
            got *= c0
            want = fd.dir_state(lambda eps: configs.stratified(rp, **{control: c0 * (1.0 + eps)})[0], EPS)
        print(f"  {control}: tangent-linear {got:.12e} fd {want:.12e}; loss {loss:.12e} oracle {fd.base:.12e}",
              flush=True)
        assert abs(loss - fd.base) <= 1e-12 * abs(fd.base), (loss, fd.base)
        assert abs(got - want) <= TOL * abs(want), (got, want)
    finally:
        fd.close()
        e.close()


def _tl_runs(rp, w, pos, d, nsteps, i_line):
    """(loss, derivative, final state, final tangent) of the FP64 and the
    mixed-precision tangent-linear runs in direction d."""
    from smaug_amd.engine import Engine

    out = []
    for prec in (P.PREC_FP64, P.PREC_MIXED):
        rp.precision = prec
        e = Engine(rp, w, pos)
        try:
            loss, dl = tl_directional(e, rp, w[3 + 2 * rp.ndim], d, None, nsteps, i_line)
            e.tl_begin(d)
            e.tl_step(nsteps)
            wv, dw = e.tl_get()
            e.tl_end()
        finally:
            e.close()
        out.append((loss, dl, wv, dw))
    rp.precision = P.PREC_FP64
    return out


def _rel(a, b):
    return np.sqrt(((a - b) ** 2).sum(axis=1) / (b ** 2).sum(axis=1))


@pytest.mark.parametrize("ni,i_line", [((16, 12), 7), ((12, 8, 8), 6), ((30, 17, 37), 9)])
def test_tl_mixed_precision_against_fp64(ni, i_line):
    """The mixed-precision tangent-linear run (FP32 value and tangent pairs in
    the stencils, FP64 dual states) against the FP64 one over 10 RK4 steps:
    states and tangents per variable and the line loss's derivative within
    1e-5 relative, and not identical (the FP32 pairs are in use)."""
    rp, w, pos = strat_tl_case(ni, len(ni), 1)
    k = P.npert(rp.ndim)
    inner = interior_mask(rp)
    rng = np.random.default_rng(13)
    scale = np.sqrt((w[:k, inner] ** 2).mean(axis=1))
    d = rng.standard_normal((k, rp.ncells)) * scale[:, None] * inner[None, :]
    (l64, d64, w64, t64), (lm, dm, wm, tm) = _tl_runs(rp, w, pos, d, 10, i_line)
    ev, et = _rel(wm, w64), _rel(tm, t64)
    print(f"[tl mixed] {ni}: states {ev.max():.2e}, tangents {et.max():.2e}, derivative {abs(dm - d64) / abs(d64):.2e}",
          flush=True)
    assert ev.max() <= 1e-5 and et.max() <= 1e-5, (ev, et)
    assert abs(dm - d64) <= 1e-5 * abs(d64), (dm, d64)
    assert et.max() > 0


def test_tl_mixed_c5_grid_192cubed():
    """The 3D adjoint configuration (192^3, FP32 shadows over an FP64 state):
    the derivative of the summed v_z^2 on the plane z ~ 4 Mm over 5 RK4 steps
    in a random direction of the initial state, from the mixed-precision
    tangent-linear run, against the FP64 tangent-linear reference (1e-5)."""
    from smaug_amd import configs

    rp = configs.stratified_params(ni=(192, 192, 192), rkon=1)
    w, pos = configs.stratified(rp, perturb=20261202)
    k = P.npert(3)
    inner = interior_mask(rp)
    rng = np.random.default_rng(17)
    scale = np.sqrt((w[:k, inner] ** 2).mean(axis=1))
    d = rng.standard_normal((k, rp.ncells)) * scale[:, None] * inner[None, :]
    i_line = 2 + int(round((4.0e6 - 0.1e6) / (5.9e6 / 192)))
    (l64, d64, _, t64), (lm, dm, _, tm) = _tl_runs(rp, w, pos, d, 5, i_line)
    et = _rel(tm, t64)
    print(f"[tl mixed] 192^3: loss {abs(lm - l64) / abs(l64):.2e}, derivative {abs(dm - d64) / abs(d64):.2e}, "
          f"tangents {et.max():.2e}", flush=True)
    assert abs(dm - d64) <= 1e-5 * abs(d64), (dm, d64)
    assert et.max() <= 1e-5, et

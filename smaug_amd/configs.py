"""Initial states for the benchmark configurations, built in FP64.

Layout of every state (as in the reference): field-major, dim 0 fastest,
ghost width 2, ``w[f, k, j, i]`` flattened to ``f*N + (k*n1 + j)*n0 + i``.
The returned arrays are numpy arrays of shape (nvar, N) and (ndim, N).

* ``orszag_tang`` -- the 2D Orszag-Tang vortex on periodic [0,1]^2 with zero
  background fields (reference model ``ot``: models/iosmaugparams.h.ot).
* ``stratified`` -- gravity-stratified solar atmosphere in hydrostatic
  balance with a uniform vertical field (reference model ``fluxtube1``:
  models/iosmaugparams.h.fluxtube1, boundary contcd4, gravity along dim 0).

Units follow the reference's conventions: SI, magnetic field in code units
B/sqrt(mu0) (magfield/B_field_3D.pro:217,349-351), gas constant R=8.31e3 and
mean molecular weight 1.2 (magfield/make_ini_file_3D.pro:96-97).
"""
from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np

from .params import (BC_CONTCD4, BC_PERIODIC, DRIVER_SPICULE1, DRIVER_TORSIONAL, DRIVER_VERTICAL, NDRV, NGHOST,
                     RunParams, nvar)

MU0 = 4.0e-7 * math.pi
R_GAS = 8.31e3
MU_MOL = 1.2


def _cell_centres(lo: float, hi: float, ni: int) -> np.ndarray:
    """Cell-centre coordinates including two ghost cells per side."""
    dx = (hi - lo) / ni
    idx = np.arange(ni + 2 * NGHOST, dtype=np.float64)
    return lo + (idx - NGHOST + 0.5) * dx


def _mesh(rp: RunParams):
    axes = [_cell_centres(rp.xmin[d], rp.xmax[d], rp.ni[d]) for d in range(rp.ndim)]
    # index order k, j, i -> arrays shaped (n2, n1, n0) so .ravel() is i-fastest
    grids = np.meshgrid(*axes[::-1], indexing="ij")[::-1]
    return [g.ravel() for g in grids]


# --------------------------------------------------------------------------
# Orszag-Tang (configuration 1)
# --------------------------------------------------------------------------

def orszag_tang_params(ni: int = 128, rkon: int = 1, courant: float = 0.5) -> RunParams:
    return RunParams(
        ndim=2, ni=[ni, ni], xmin=[0.0, 0.0], xmax=[1.0, 1.0],
        dt=1.0, gamma=5.0 / 3.0, courant=courant, rkon=rkon, moddton=1,
        hyperdifmom=1, chyp={"default": 0.2},
        bcu=[BC_PERIODIC] * 3, bcl=[BC_PERIODIC] * 3)


def orszag_tang(rp: RunParams) -> Tuple[np.ndarray, np.ndarray]:
    """rho=25/(36 pi), p=5/(12 pi), v=(-sin 2pi y, sin 2pi x),
    B=B0(-sin 2pi y, sin 4pi x), B0=1/sqrt(4 pi); background zero."""
    x, y = _mesh(rp)
    gam = rp.gamma
    rho = np.full_like(x, 25.0 / (36.0 * math.pi))
    pr = 5.0 / (12.0 * math.pi)
    vx = -np.sin(2.0 * math.pi * y)
    vy = np.sin(2.0 * math.pi * x)
    b0 = 1.0 / math.sqrt(4.0 * math.pi)
    bx = -b0 * np.sin(2.0 * math.pi * y)
    by = b0 * np.sin(4.0 * math.pi * x)
    e = pr / (gam - 1.0) + 0.5 * rho * (vx * vx + vy * vy) + 0.5 * (bx * bx + by * by)
    w = np.zeros((nvar(2), x.size))
    w[0], w[1], w[2], w[3], w[4], w[5] = rho, rho * vx, rho * vy, e, bx, by
    pos = np.stack([x, y])
    return w, pos


# --------------------------------------------------------------------------
# Stratified atmosphere (configurations 2-5)
# --------------------------------------------------------------------------

def temperature(z: np.ndarray, t_lo=6000.0, t_hi=1.0e6, z_tr=2.0e6, w_tr=0.2e6) -> np.ndarray:
    """T(z) = t_lo + (t_hi - t_lo)(1 + tanh((z - z_tr)/w_tr))/2."""
    return t_lo + (t_hi - t_lo) * (1.0 + np.tanh((z - z_tr) / w_tr)) / 2.0


def _inv_t_integral(z, t_lo, t_hi, z_tr, w_tr):
    """Antiderivative of 1/T(z): with T = a + b tanh(u), u=(z-z_tr)/w,
    int du/(a + b tanh u) = (a u - b ln(a cosh u + b sinh u))/(a^2 - b^2)."""
    a = 0.5 * (t_lo + t_hi)
    b = 0.5 * (t_hi - t_lo)
    u = (np.asarray(z, dtype=np.float64) - z_tr) / w_tr
    # ln(a cosh u + b sinh u) = ln(((a+b)e^u + (a-b)e^-u)/2), evaluated stably
    lg = np.logaddexp(math.log(a + b) + u, math.log(a - b) - u) - math.log(2.0)
    return w_tr * (a * u - b * lg) / (a * a - b * b)


def hydrostatic(z: np.ndarray, rho0=2.7e-4, z0=0.1e6, grav=274.0, t_lo=6000.0, t_hi=1.0e6,
                z_tr=2.0e6, w_tr=0.2e6) -> Tuple[np.ndarray, np.ndarray]:
    """Pressure and density of dp/dz = -rho g with rho = p mu/(R T), integrated
    in closed form: p = p0 exp(-(mu g/R) int_{z0}^{z} dz/T)."""
    t0 = temperature(np.array([z0]), t_lo, t_hi, z_tr, w_tr)[0]
    p0 = rho0 * R_GAS * t0 / MU_MOL
    iz = _inv_t_integral(z, t_lo, t_hi, z_tr, w_tr) - _inv_t_integral(z0, t_lo, t_hi, z_tr, w_tr)
    p = p0 * np.exp(-(MU_MOL * grav / R_GAS) * iz)
    rho = p * MU_MOL / (R_GAS * temperature(z, t_lo, t_hi, z_tr, w_tr))
    return p, rho


def stratified_dt(ni, ndim: int, courant: float, b_tesla: float = 0.01, gamma: float = 5.0 / 3.0) -> float:
    """Fixed time step of the stratified runs: courant * min(dx) over the
    largest fast speed sqrt((gamma p_b + B_b^2)/rho_b) of the background,
    rounded down to 3 significant digits."""
    xmin = [0.1e6, 0.0, 0.0][:ndim]
    xmax = [6.0e6, 4.0e6, 4.0e6][:ndim]
    z = _cell_centres(xmin[0], xmax[0], ni[0])
    p, rho = hydrostatic(z)
    bz = b_tesla / math.sqrt(MU0)
    cmax = float(np.max(np.sqrt((gamma * p + bz * bz) / rho)))
    dx = min((xmax[d] - xmin[d]) / ni[d] for d in range(ndim))
    dt = courant * dx / cmax
    e = math.floor(math.log10(dt)) - 2
    return float(f"{math.floor(dt / 10.0 ** e)}e{e}")


def stratified_params(ni=(128, 128, 128), ndim: int = 3, rkon: int = 1, courant: float = 0.15,
                      b_tesla: float = 0.01, amplitude: float = 200.0) -> RunParams:
    """z (dim 0) in [0.1, 6] Mm, x, y (dims 1, 2) in [0, 4] Mm, gravity -274 m/s^2
    along dim 0, contcd4 boundaries, hyperdiffusion 0.02 (rho, e, b) and 0.2 (m).
    Fixed time step (PMODDTON 0, as models/iosmaugparams.h.fluxtube1:27), sized
    by stratified_dt so the run is stable at the chosen resolution."""
    ni = list(ni)[:ndim]
    xmin = [0.1e6, 0.0, 0.0][:ndim]
    xmax = [6.0e6, 4.0e6, 4.0e6][:ndim]
    drv = [amplitude, 300.0, 0.1e6, 2.0e6, 2.0e6, 0.05e6, 0.1e6] + [0.0] * (NDRV - 7)
    return RunParams(
        ndim=ndim, ni=ni, xmin=xmin, xmax=xmax, dt=stratified_dt(ni, ndim, courant, b_tesla),
        gamma=5.0 / 3.0, g=[-274.0, 0.0, 0.0], courant=courant, rkon=rkon, moddton=0, hyperdifmom=1,
        chyp={"default": 0.02, "rho": 0.02, "energy": 0.02, "b1": 0.02, "b2": 0.02, "b3": 0.02,
              "mom1": 0.2, "mom2": 0.2, "mom3": 0.2},
        bcu=[BC_CONTCD4] * 3, bcl=[BC_CONTCD4] * 3, driver=DRIVER_VERTICAL, drv=drv)


def fluxtube1_driver(rp: RunParams, aa: float = 10.0, tdep: float = 1.0, centre=(0.1e6, 1.0e6, 1.0e6),
                     widths=(0.05e6, 0.1e6, 0.1e6), scales=(2.0e6, 2.0e6)) -> RunParams:
    """The torsional driver of the fluxtube1 model on rp (3D): the constants of
    models/usersource.cu.fluxtube1:28-54 by default (aa 10 m/s, tdep 1,
    centre (0.1, 1, 1) Mm along dims 0, 1, 2, Gaussian widths 0.05, 0.1,
    0.1 Mm, xxmax = yymax = 2 Mm)."""
    rp.driver = DRIVER_TORSIONAL
    rp.drv = [aa, tdep, *centre, *widths, *scales] + [0.0] * (NDRV - 10)
    return rp


def spicule1_driver(rp: RunParams, AA: float = 200.0, tdep: float = 1.0, centre=(0.0174e6, 1.998046e6),
                    widths=(0.001e6, 0.001e6)) -> RunParams:
    """The spicule1 model's driver on rp: the constants of
    models/usersource.cu.spicule1:32-65 by default (AA 200 m/s, tdep 1,
    centre (0.0174, 1.998046) Mm along dims 0, 1, widths delta_z =
    delta_x = 0.001 Mm); it acts in the step whose counter is 1 only."""
    rp.driver = DRIVER_SPICULE1
    rp.drv = [AA, tdep, centre[0], centre[1], 0.0, widths[0], widths[1]] + [0.0] * (NDRV - 7)
    return rp


def stratified(rp: RunParams, b_tesla: float = 0.01, perturb: Optional[int] = None,
               amp: float = 1.0e-4, t_photo: float = 6000.0, z_tr: float = 2.0e6) -> Tuple[np.ndarray, np.ndarray]:
    """Hydrostatic background with a uniform vertical field; perturbation
    zero, or seeded noise (seed `perturb`) smoothed by three passes of a
    3-point box filter: drho/rho_b ~ N(0, amp), de ~ N(0, amp) p/(gamma-1),
    v ~ N(0, amp c_s), db ~ N(0, amp |b_b|).  t_photo and z_tr are the
    photospheric temperature and the transition height of T(z)."""
    coords = _mesh(rp)
    n0 = rp.n[0]
    gam = rp.gamma
    # the background depends on dim 0 (height) only: build it on the axis
    zax = coords[0][:n0]
    p_ax, rho_ax = hydrostatic(zax, grav=-rp.g[0] if rp.g[0] != 0 else 274.0, t_lo=t_photo, z_tr=z_tr)
    reps = rp.ncells // n0
    p = np.tile(p_ax, reps)
    rho_b = np.tile(rho_ax, reps)
    bz = b_tesla / math.sqrt(MU0)
    nd = rp.ndim
    ncell = rp.ncells
    w = np.zeros((nvar(nd), ncell))
    eb = p / (gam - 1.0) + 0.5 * bz * bz
    i_eb, i_rhob, i_bb = 2 + 2 * nd, 3 + 2 * nd, 4 + 2 * nd
    w[i_eb] = eb
    w[i_rhob] = rho_b
    w[i_bb] = bz
    if perturb is not None:
        rng = np.random.default_rng(perturb)
        shape = tuple(rp.n[::-1])

        taps = np.array([1.0, 3.0, 6.0, 7.0, 6.0, 3.0, 1.0]) / 27.0

        def smooth(a):
            # three passes of the periodic 3-point box filter along every axis,
            # applied as the equivalent 7-tap kernel [1,3,6,7,6,3,1]/27
            a = a.reshape(shape).astype(np.float32)
            for ax in range(a.ndim):
                out = np.zeros_like(a)
                for t, wgt in zip(range(-3, 4), taps):
                    out += np.float32(wgt) * np.roll(a, t, axis=ax)
                a = out
            return a.ravel().astype(np.float64)

        cs = np.sqrt(gam * p / rho_b)
        drho = smooth(rng.normal(0.0, amp, ncell)) * rho_b
        w[0] = drho
        for d in range(nd):
            v = smooth(rng.normal(0.0, amp, ncell)) * cs
            w[1 + d] = (rho_b + drho) * v
        de = smooth(rng.normal(0.0, amp, ncell)) * p / (gam - 1.0)
        w[1 + nd] = de
        for d in range(nd):
            w[2 + nd + d] = smooth(rng.normal(0.0, amp, ncell)) * bz
    pos = np.stack(coords)
    return w, pos


STATE_CONTROLS = ("t_photo", "z_tr")


def stratified_tangent(rp: RunParams, control: str, rel: float = 1.0e-4, **kw) -> np.ndarray:
    """d(state)/d(control) of the stratified initial state (nvar x N: the
    perturbation, which the seeded noise scales with the background, and the
    background fields), for control 't_photo' or 'z_tr' at the values in kw
    (defaults 6000 K, 2 Mm).  Richardson-extrapolated central differences of
    the closed-form builder at steps rel and 2 rel of the control: truncation
    O(rel^4), roundoff ~1e-16/rel, ~1e-11 relative at the default.  The
    background direction for smaug_tl_begin (dbg) is rows np.., the evolved
    direction (dw0) rows ..np."""
    if control not in STATE_CONTROLS:
        raise ValueError(f"control must be one of {STATE_CONTROLS}")
    base = dict(t_photo=6000.0, z_tr=2.0e6)
    base.update(kw)
    c0 = base[control]

    def at(h):
        args = dict(base)
        args[control] = c0 + h
        return stratified(rp, **args)[0]

    h = rel * abs(c0)

    def central(hh):
        return (at(hh) - at(-hh)) / (2.0 * hh)

    return (4.0 * central(h) - central(2.0 * h)) / 3.0

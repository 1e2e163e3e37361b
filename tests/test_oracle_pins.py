"""Known-answer and invariant tests that pin the CPU oracle (no GPU).

The reference holds no result vectors (it has no tests), so the oracle is
pinned by what the mathematics fixes: fixed points, polynomial exactness of
the difference operators, conservation on periodic grids, the convergence
order of the four-step RK scheme, textbook wave speeds, brute-force windows
of the hyperdiffusion coefficient, and closed forms of the gravity and
hyperdiffusion updates.  Each test names the reference code it pins.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle.pyoracle import Oracle
from smaug_amd import configs
from smaug_amd.params import (BC_CONTCD4, BC_PERIODIC, T_ETERM, T_FLUX, T_GRAV, T_HE, T_HRHO, T_SOURCE, TERMS_ALL,
                              RunParams, npert, nvar)

from cases import ot_case, strat_case


# ---------------------------------------------------------------- helpers

def box_params(ndim=2, ni=(12, 10, 8), bc=BC_PERIODIC, terms=TERMS_ALL, rkon=0, dt=1.0e-3,
               chyp=0.02, hyper=1, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), g=(0.0, 0.0, 0.0)):
    """A fixed-dt run on a uniform box (moddton=0 so dt = DT every step)."""
    return RunParams(ndim=ndim, ni=list(ni[:ndim]), xmin=list(lo[:ndim]), xmax=list(hi[:ndim]),
                     dt=dt, gamma=5.0 / 3.0, g=list(g), courant=0.3, rkon=rkon, moddton=0,
                     hyperdifmom=hyper, chyp={"default": chyp}, bcu=[bc] * 3, bcl=[bc] * 3,
                     terms=terms)


def mesh(rp):
    """Cell-centre coordinates (ndim, N), i fastest, ghosts included."""
    return np.stack(configs._mesh(rp))


def uniform_background(rp, rho_b=1.0, p_b=0.6, b_b=(0.0, 0.0, 0.0)):
    nd = rp.ndim
    w = np.zeros((nvar(nd), rp.ncells))
    w[2 + 2 * nd] = p_b / (rp.gamma - 1.0) + 0.5 * sum(b * b for b in b_b[:nd])
    w[3 + 2 * nd] = rho_b
    for d in range(nd):
        w[4 + 2 * nd + d] = b_b[d]
    return w


def grid_view(rp, a):
    """(.., N) -> (.., n2, n1, n0) view (k, j, i)."""
    return a.reshape(a.shape[:-1] + tuple(rp.n[::-1]))


def core_mask(rp, margin):
    """Cells at least `margin` cells from every grid edge."""
    m = np.ones(tuple(rp.n[::-1]), dtype=bool)
    for d in range(rp.ndim):
        ax = rp.ndim - 1 - d
        idx = np.arange(rp.n[d])
        keep = (idx >= margin) & (idx <= rp.n[d] - 1 - margin)
        shape = [1] * rp.ndim
        shape[ax] = rp.n[d]
        m &= keep.reshape(shape)
    return m.ravel()


def one_step(rp, w, pos, n=1):
    o = Oracle(rp, w, pos)
    o.step(n)
    out = o.state()
    o.close()
    return out


# ---------------------------------------------------------------- fixed points

@pytest.mark.parametrize("rkon", [0, 1])
def test_zero_perturbation_is_a_fixed_point(rkon):
    """w' = 0 over any background: every flux, source and hyperdiffusion
    term of the split equations vanishes (src/centdiff1.cu:51-219,
    src/centdiff2.cu:80-356, src/hyperdifvisc1r.cu:426-565 takes the
    zero-denominator branch), so the state stays exactly zero."""
    rp, w, pos = strat_case(ni=(12, 10, 10), rkon=rkon, driver=False)
    w[: npert(3)] = 0.0
    out = one_step(rp, w, pos, n=2)
    assert np.all(out[: npert(3)] == 0.0)
    assert np.array_equal(out[npert(3):], w[npert(3):])


@pytest.mark.parametrize("rkon", [0, 1])
def test_uniform_state_is_a_fixed_point(rkon):
    """A uniform perturbation on a periodic grid: constant fluxes have zero
    divergence and constant q gives nu = 0, bitwise."""
    rp = box_params(ndim=2, rkon=rkon)
    w = uniform_background(rp, rho_b=0.0, p_b=0.0)
    w[0], w[1], w[2], w[3], w[4], w[5] = 1.3, 0.2, -0.1, 2.0, 0.3, 0.4
    out = one_step(rp, w, mesh(rp), n=2)
    assert np.array_equal(out, w)


# ---------------------------------------------------------------- stencils

def _poly_state(rp):
    o = Oracle(rp, uniform_background(rp), mesh(rp))
    return o


@pytest.mark.parametrize("ndim,direction", [(2, 0), (2, 1), (3, 0), (3, 1), (3, 2)])
def test_grad3dn_exact_on_quartics(ndim, direction):
    """grad3dn (src/gradops.cu:437-602) is the 4th-order central difference:
    exact for polynomials of degree <= 4 on a uniform grid."""
    rp = box_params(ndim=ndim, ni=(12, 10, 8), bc=BC_PERIODIC if ndim == 2 else BC_CONTCD4)
    X = mesh(rp)
    x = X[direction]
    other = X[(direction + 1) % ndim]
    a = (0.3 - 1.1 * x + 0.7 * x ** 2 - 2.0 * x ** 3 + 1.5 * x ** 4) * (1.0 + other ** 2)
    da = (-1.1 + 1.4 * x - 6.0 * x ** 2 + 6.0 * x ** 3) * (1.0 + other ** 2)
    o = _poly_state(rp)
    # bcidx 0 row is periodic in 2D; in 3D pick the cells the Q1 zeroing spares
    g = o.grad("grad3dn", a, 0, direction)
    o.close()
    m = core_mask(rp, 4) if ndim == 3 else core_mask(rp, 2)
    assert np.allclose(g[m], da[m], rtol=1e-11, atol=1e-11)
    # a cubic in the other direction has no derivative along this one
    b = (1.0 + other ** 3)
    o = _poly_state(rp)
    gb = o.grad("grad3dn", b, 0, direction)
    o.close()
    assert np.all(np.abs(gb[m]) < 1e-13)


def test_grad3dn_is_not_exact_beyond_degree_four():
    """Sanity of the test above: a quintic leaves the 4th-order error term."""
    rp = box_params(ndim=2)
    X = mesh(rp)
    a = X[0] ** 5
    o = _poly_state(rp)
    g = o.grad("grad3dn", a, 0, 0)
    o.close()
    m = core_mask(rp, 2)
    h = rp.dx[0]
    # truncation error of the 5-point first derivative: -h^4/30 f^(5) = -4 h^4
    assert np.allclose(g[m] - 5 * X[0][m] ** 4, -4.0 * h ** 4, rtol=1e-6)


@pytest.mark.parametrize("direction", [0, 1])
def test_one_sided_and_central_differences(direction):
    """grad1r3n / grad1l3n (src/gradops.cu:608-750) are the forward and
    backward first differences, grad13n (:755-823) the central one: on a
    quadratic they give f'(x + h/2), f'(x - h/2) and f'(x) exactly."""
    rp = box_params(ndim=2)
    X = mesh(rp)
    x = X[direction]
    h = rp.dx[direction]
    a = 2.0 * x ** 2 - 3.0 * x + 0.25
    fp = lambda z: 4.0 * z - 3.0  # noqa: E731
    o = _poly_state(rp)
    gr, gl, gc = (o.grad(k, a, 0, direction) for k in ("grad1r3n", "grad1l3n", "grad13n"))
    o.close()
    m = core_mask(rp, 1)
    assert np.allclose(gr[m], fp(x[m] + h / 2), rtol=1e-12, atol=1e-12)
    assert np.allclose(gl[m], fp(x[m] - h / 2), rtol=1e-12, atol=1e-12)
    assert np.allclose(gc[m], fp(x[m]), rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- nu

def _visc_case(q_fn, ndim=2, field=0):
    rp = box_params(ndim=ndim, ni=(16, 12, 10), bc=BC_PERIODIC if ndim == 2 else BC_CONTCD4, chyp=0.05)
    X = mesh(rp)
    w = uniform_background(rp, rho_b=1.0, p_b=0.6, b_b=(0.2, 0.1, 0.0))
    w[field] = q_fn(X)
    return rp, w, X


def _nu_brute(q, dxk, s, n, margin):
    """nu at i+1/2 and i-1/2 from the textbook windows: max |3rd difference|
    over the 3 interfaces around the face, over max |1st difference| over
    the 5 interfaces around it (Caunt & Korpi 2001, as in SAC)."""
    def d3(a, i):  # third difference at interface i+1/2
        return abs(a[i + 2] - 3 * a[i + 1] + 3 * a[i] - a[i - 1])

    def d1(a, i):
        return abs(a[i + 1] - a[i])

    nur, nul = np.zeros(n), np.zeros(n)
    for i in range(margin, n - margin):
        for face, out in ((i, nur), (i - 1, nul)):
            num = max(d3(q, face + m) for m in (-1, 0, 1))
            den = max(d1(q, face + m) for m in (-2, -1, 0, 1, 2))
            out[i] = dxk * num / den if den > 0 else 0.0
    return nur, nul


@pytest.mark.parametrize("dim", [0, 1])
def test_nu_matches_brute_force_windows(dim):
    """hyperdifvisc1r/1l chains (src/hyperdifvisc1r.cu:426-843,
    src/hyperdifvisc1l.cu:723-853) against a brute-force window maximum on
    every line of a random smooth field, away from the halo quirks."""
    rng = np.random.default_rng(7)
    rp, w, X = _visc_case(lambda X: 0.0)
    wv = grid_view(rp, w)
    wv[0] = 0.05 * rng.standard_normal(wv[0].shape)
    o = Oracle(rp, w, mesh(rp))
    nur, nul, cmax = o.visc(0, dim)
    rho = grid_view(rp, o.state()[0])  # ghosts hold the periodic images
    o.close()
    assert cmax > 0
    dxk = rp.dx[dim] * cmax * 0.05
    NR, NL = grid_view(rp, nur), grid_view(rp, nul)
    margin = 4
    n1, n0 = rp.n[1], rp.n[0]
    checked = 0
    if dim == 0:
        for j in range(margin, n1 - margin):
            br, bl = _nu_brute(rho[j, :], dxk, 1, n0, margin)
            sl = slice(margin, n0 - margin)
            assert np.allclose(NR[j, sl], br[sl], rtol=1e-12, atol=0)
            assert np.allclose(NL[j, sl], bl[sl], rtol=1e-12, atol=0)
            checked += 1
    else:
        for i in range(margin, n0 - margin):
            br, bl = _nu_brute(rho[:, i], dxk, 1, n1, margin)
            sl = slice(margin, n1 - margin)
            assert np.allclose(NR[sl, i], br[sl], rtol=1e-12, atol=0)
            assert np.allclose(NL[sl, i], bl[sl], rtol=1e-12, atol=0)
            checked += 1
    assert checked > 0
    assert np.all(nur >= 0) and np.all(nul >= 0)


def test_nu_left_face_is_right_face_of_neighbour():
    """nu_{i-1/2} at cell i and nu_{i+1/2} at cell i-1 use the same window."""
    rng = np.random.default_rng(11)
    rp, w, X = _visc_case(lambda X: 0.0)
    w[0] = 0.05 * rng.standard_normal(rp.ncells)
    o = Oracle(rp, w, mesh(rp))
    nur, nul, _ = o.visc(0, 0)
    o.close()
    NR, NL = grid_view(rp, nur), grid_view(rp, nul)
    assert np.array_equal(NL[4:-4, 5:-4], NR[4:-4, 4:-5])


def test_nu_vanishes_for_linear_and_constant_q():
    """Linear q: every third difference is 0 (integer-valued, so exactly);
    constant q: the zero-denominator branch gives nu = 0 on every cell."""
    rp, w, X = _visc_case(lambda X: 0.0)
    wv = grid_view(rp, w)
    jj, ii = np.meshgrid(np.arange(rp.n[1]), np.arange(rp.n[0]), indexing="ij")
    wv[0] = (2.0 * ii + 3.0 * jj) * 2.0 ** -10
    o = Oracle(rp, w, mesh(rp))
    for dim in (0, 1):
        nur, nul, _ = o.visc(0, dim)
        m = core_mask(rp, 5)  # windows inside the interior (ghosts hold periodic images)
        assert np.all(nur[m] == 0.0) and np.all(nul[m] == 0.0)
    o.close()
    w[0] = 0.01
    o = Oracle(rp, w, mesh(rp))
    for dim in (0, 1):
        nur, nul, _ = o.visc(0, dim)
        assert np.all(nul == 0.0)
        # the right-chain window of the last interior layer reaches the
        # padded halo that only the stray write fills (quirk Q7)
        last = np.moveaxis(grid_view(rp, nur), 1 - dim, 0)
        assert np.all(np.delete(last, rp.n[dim] - 3, axis=0) == 0.0)
    o.close()


# ---------------------------------------------------------------- fluxes

@pytest.mark.parametrize("ndim,direction", [(2, 0), (2, 1), (3, 2)])
def test_mass_flux_divergence_exact_on_cubic_momentum(ndim, direction):
    """rho_t = 1 so the mass flux is m_d itself (src/centdiff1.cu:51-60);
    one Euler step with the flux term only gives rho = -dt d(m_d)/dx_d
    exactly for a cubic m_d (4th-order divergence, src/centdiff1.cu:19-43)."""
    dt = 1.0e-3
    rp = box_params(ndim=ndim, ni=(12, 12, 12), bc=BC_CONTCD4, terms=T_FLUX, dt=dt, hyper=0)
    X = mesh(rp)
    x = X[direction]
    w = uniform_background(rp, rho_b=1.0, p_b=0.6)
    eps = 1e-3
    w[1 + direction] = eps * (0.5 + x - 2.0 * x ** 2 + 3.0 * x ** 3)
    out = one_step(rp, w, X)
    expect = -dt * eps * (1.0 - 4.0 * x + 9.0 * x ** 2)
    m = core_mask(rp, 4)
    assert np.allclose(out[0][m], expect[m], rtol=1e-9, atol=1e-18)


@pytest.mark.parametrize("hyper", [0, 1])
def test_periodic_flux_form_conserves(hyper):
    """Orszag-Tang on the periodic grid: the divergence of the fluxes
    telescopes, so the interior sums of rho, m, e and b are conserved by an
    Euler step (hyperdiffusion of rho and e is in flux form too)."""
    rp, w, pos = ot_case(ni=16, rkon=0)
    rp.moddton = 0
    rp.dt = 2.0e-3
    rp.hyperdifmom = hyper
    rp.terms = T_FLUX | (T_HRHO | T_HE if hyper else 0)
    out = one_step(rp, w, pos)
    m = core_mask(rp, 2)
    for f in range(npert(2)):
        before, after = w[f][m].sum(), out[f][m].sum()
        scale = np.abs(w[f][m]).sum() + 1e-30
        assert abs(after - before) <= 1e-13 * scale, (f, after - before, scale)


# ---------------------------------------------------------------- time scheme

def _wave_run(nsteps, T, rkon=1):
    rp = box_params(ndim=2, ni=(16, 8), bc=BC_PERIODIC, terms=T_FLUX, rkon=rkon, dt=T / nsteps,
                    hyper=0)
    X = mesh(rp)
    w = uniform_background(rp, rho_b=1.0, p_b=0.6)
    k = 2 * math.pi
    eps = 1e-3
    w[0] = eps * np.sin(k * X[0])
    w[1] = eps * 0.9 * np.sin(k * X[0])
    w[3] = eps * 0.9 / (rp.gamma - 1.0) * np.sin(k * X[0])
    return one_step(rp, w, X, n=nsteps)


def test_rk4_is_fourth_order_in_time():
    """The four-step scheme (src/advance.cu:76-99 on top of the stage
    operator) converges at 4th order: halving dt cuts the time error ~16x
    (against a 64x finer run on the same grid)."""
    T = 0.2
    ref = _wave_run(256, T)
    e1 = np.abs(_wave_run(8, T) - ref)[:4].max()
    e2 = np.abs(_wave_run(16, T) - ref)[:4].max()
    order = math.log2(e1 / e2)
    assert 3.6 < order < 4.6, order


def test_euler_is_first_order_in_time():
    T = 0.2
    ref = _wave_run(512, T, rkon=0)
    e1 = np.abs(_wave_run(16, T, rkon=0) - ref)[:4].max()
    e2 = np.abs(_wave_run(32, T, rkon=0) - ref)[:4].max()
    order = math.log2(e1 / e2)
    assert 0.8 < order < 1.3, order


# ---------------------------------------------------------------- boundaries, grid

def test_periodic_ghost_images():
    """Periodic fill (src/boundary.cu:72-225): ghost g <- g + (n-4) below,
    g <- g - (n-4) above, on every evolved field after a step."""
    rp, w, pos = ot_case(ni=12, rkon=0)
    out = grid_view(rp, one_step(rp, w, pos)[: npert(2)])
    n0, n1 = rp.n[0], rp.n[1]
    for g, s in ((0, n0 - 4), (1, n0 - 3), (n0 - 2, 2), (n0 - 1, 3)):
        assert np.array_equal(out[:, :, g], out[:, :, s])
    for g, s in ((0, n1 - 4), (1, n1 - 3), (n1 - 2, 2), (n1 - 1, 3)):
        assert np.array_equal(out[:, g, :], out[:, s, :])


def test_contcd4_ghost_images():
    """contcd4 fill (models/boundary.cu.fluxtube1:16-85): 0 <- 4, 1 <- 3,
    n-1 <- n-5, n-2 <- n-4 along each direction."""
    rp, w, pos = strat_case(ni=(10, 8, 8), rkon=0)
    out = grid_view(rp, one_step(rp, w, pos)[: npert(3)])
    n = rp.n
    for g, s in ((0, 4), (1, 3), (n[0] - 1, n[0] - 5), (n[0] - 2, n[0] - 4)):
        assert np.array_equal(out[..., g], out[..., s])
    for g, s in ((0, 4), (1, 3), (n[2] - 1, n[2] - 5), (n[2] - 2, n[2] - 4)):
        assert np.array_equal(out[:, g], out[:, s])


@pytest.mark.parametrize("ndim", [2, 3])
def test_uniform_grid_spacing(ndim):
    """initgrid (src/init.cu:2502-3051): on a uniform grid delx = dx on every
    cell (ghosts extrapolated linearly), positions unchanged inside."""
    rp = box_params(ndim=ndim, ni=(10, 9, 8), bc=BC_PERIODIC if ndim == 2 else BC_CONTCD4,
                    lo=(0.5, -1.0, 2.0), hi=(1.5, 3.0, 2.5))
    X = mesh(rp)
    o = Oracle(rp, uniform_background(rp), X)
    pos, delx = o.grid()
    o.close()
    for d in range(ndim):
        assert np.allclose(delx[d], rp.dx[d], rtol=1e-12, atol=0)
    m = core_mask(rp, 2)
    assert np.array_equal(pos[:, m], X[:, m])


# ---------------------------------------------------------------- wave speeds, dt

def test_cfl_matches_textbook_fast_speed():
    """cucomputec/cucomputemaxc (src/dervfields.cu:241-309,
    src/computedervfields.cu:1864-2050): cmax per direction is the maximum
    over all cells of the fast magnetosonic speed plus |v_d|, with the
    thermal pressure of the total state; the CFL dt uses the last
    direction's cmax (src/iosmaug.c:541-616)."""
    rp, w, pos = strat_case(ni=(10, 8, 8), rkon=0)
    o = Oracle(rp, w, pos)
    cm, dt = o.cfl()
    w = o.state()  # ghosts filled by the boundary conditions
    o.close()
    nd, gam = 3, rp.gamma
    rho = w[0] + w[3 + 2 * nd]
    m = w[1:1 + nd]
    B = w[2 + nd:2 + 2 * nd] + w[4 + 2 * nd:4 + 3 * nd]
    etot = w[1 + nd] + w[2 + 2 * nd]
    b2 = (w[2 + nd:2 + 2 * nd] ** 2).sum(0)
    bbb = (w[2 + nd:2 + 2 * nd] * w[4 + 2 * nd:4 + 3 * nd]).sum(0)
    bb2 = (w[4 + 2 * nd:4 + 3 * nd] ** 2).sum(0)
    # SAC energy split: e_b holds p_b/(gamma-1) + b_b^2/2, e the rest
    p = (gam - 1) * (etot - 0.5 * (m ** 2).sum(0) / rho - 0.5 * (b2 + 2 * bbb + bb2))
    cs2 = gam * p / rho
    va2 = (B ** 2).sum(0) / rho
    for d in range(nd):
        vad2 = B[d] ** 2 / rho
        cf = np.sqrt(0.5 * (cs2 + va2 + np.sqrt((cs2 + va2) ** 2 - 4 * cs2 * vad2))) + np.abs(m[d] / rho)
        assert cm[d] == pytest.approx(max(cf.max(), np.sqrt(cs2).max()), rel=1e-12)
    assert dt == pytest.approx(min(rp.dt, rp.courant * min(rp.dx) / cm[-1]), rel=1e-14)


@pytest.mark.parametrize("rkon", [0, 1])
def test_cfl_step_restarts_from_dt_each_step(rkon):
    """With moddton=1 every step starts from dt = DT (src/iosmaug.c:549-550)
    and takes the CFL step only when it is below DT (:568-569): a DT below
    the CFL limit gives dt == DT on every step and t == n*DT exactly (the
    initial 1e-8 of :333-336 is never used), a DT above it gives the CFL
    step of the state at the start of each step."""
    rp, w, pos = ot_case(ni=16, rkon=rkon)
    o = Oracle(rp, w, pos)
    _, dt_cfl = o.cfl()
    o.close()
    assert dt_cfl < rp.dt  # the OT configuration's DT (1.0) is a cap only
    small = dt_cfl / 7.0
    rp.dt = small
    o = Oracle(rp, w, pos)
    t = 0.0
    for n in range(1, 4):
        o.step(1)
        it, tt, qt, dt = o.clock()
        t += small
        assert dt == small and tt == t and qt == t, (n, dt, tt, qt)
    o.close()
    rp.dt = 1.0
    o = Oracle(rp, w, pos)
    for n in range(3):
        _, want = o.cfl()
        o.step(1)
        assert o.clock()[3] == want and want < 1.0
    o.close()


# ---------------------------------------------------------------- source terms

def test_gravity_closed_form():
    """addgrav_cd2 (src/centdiff2.cu:118-154): m_d += dt g_d rho',
    e += dt rho' g.v with v = m/rho_t, on interior cells."""
    dt = 0.5
    rp = box_params(ndim=3, ni=(8, 8, 8), bc=BC_CONTCD4, terms=T_GRAV, dt=dt, g=(-2.5, 0.75, 1.25))
    rng = np.random.default_rng(3)
    w = uniform_background(rp, rho_b=1.0, p_b=0.6)
    w[0] = 0.1 * rng.standard_normal(rp.ncells)
    for d in range(3):
        w[1 + d] = 0.1 * rng.standard_normal(rp.ncells)
    out = one_step(rp, w, mesh(rp))
    m = core_mask(rp, 2)
    rt = w[0] + w[9]
    for d in range(3):
        assert np.allclose(out[1 + d][m], w[1 + d][m] + dt * rp.g[d] * w[0][m], rtol=1e-14, atol=1e-15)
    gv = sum(rp.g[d] * w[1 + d] for d in range(3)) / rt
    assert np.allclose(out[4][m], w[4][m] + dt * w[0][m] * gv[m], rtol=1e-13, atol=1e-15)
    assert np.array_equal(out[0][m], w[0][m])


def test_driver_scaling_and_phase():
    """Vertical-velocity driver (models/usersource.cu.fluxtube1:3-90 form):
    momentum injection is linear in the amplitude and kinetic-energy
    injection quadratic (doubling A doubles dm and quadruples de, exactly);
    at driver phase 0 nothing is injected."""
    def run(A, qt0):
        rp, w, pos = strat_case(ni=(10, 8, 8), rkon=0)
        rp.terms = T_SOURCE
        rp.moddton = 0
        rp.dt = 2.0
        # a wide driver centred in the box so every cell sees it
        rp.drv = [A, 300.0, 3.0e6, 2.0e6, 2.0e6, 2.0e6, 2.0e6, 0.0]
        rp.qt0 = qt0
        w[1:5] = 0.0  # m and e start at zero: the update is the injection itself
        return one_step(rp, w, pos)

    d1 = run(100.0, 37.0)
    d2 = run(200.0, 37.0)
    m = core_mask(strat_case(ni=(10, 8, 8))[0], 2)
    assert np.abs(d1[1][m]).min() > 0 and np.abs(d1[4][m]).min() > 0
    assert np.array_equal(d2[1][m], 2 * d1[1][m])
    assert np.array_equal(d2[4][m], 4 * d1[4][m])
    assert np.all(d1[2:4][:, m] == 0.0)
    # energy injected is the kinetic energy of the injected velocity:
    # de = dt rho_t vz^2 / 2 with vz = dm / (dt rho_t)
    rp, w, _ = strat_case(ni=(10, 8, 8), rkon=0)
    rt = w[0] + w[9]
    assert np.allclose(d1[4][m], d1[1][m] ** 2 / (2.0 * 2.0 * rt[m]), rtol=1e-12, atol=0)
    # qt advances by DT before the stage: qt0 = -DT puts the phase at 0
    d0 = run(100.0, -2.0)
    assert np.all(d0[1][m] == 0.0) and np.all(d0[4][m] == 0.0)


@pytest.mark.parametrize("dim", [0, 1])
def test_rho_hyperdiffusion_closed_form(dim):
    """hyperdifrhosource1 (src/hyperdifrhosource1.cu:15-275): with the
    coefficients nu_r, nu_l of the rho field, one Euler step adds
    dt (nu_r (rho_{i+1}-rho_i) - nu_l (rho_i-rho_{i-1}))/dx^2 per direction."""
    dt = 1.0e-3
    rp = box_params(ndim=2, ni=(16, 12), bc=BC_PERIODIC, terms=T_HRHO, dt=dt, chyp=0.05)
    rng = np.random.default_rng(5)
    w = uniform_background(rp, rho_b=1.0, p_b=0.6, b_b=(0.2, 0.1, 0.0))
    w[0] = 0.05 * rng.standard_normal(rp.ncells)
    X = mesh(rp)
    o = Oracle(rp, w, X)
    state = o.state()
    nus = [o.visc(0, d) for d in range(2)]
    o.close()
    out = one_step(rp, w, X)
    rho = grid_view(rp, state[0])
    expect = rho.copy()
    for d, (nur, nul, _) in enumerate(nus):
        ax = 1 - d
        h = rp.dx[d]
        rp1, rm1 = np.roll(rho, -1, axis=ax), np.roll(rho, 1, axis=ax)
        expect = expect + dt * (grid_view(rp, nur) * (rp1 - rho) / h - grid_view(rp, nul) * (rho - rm1) / h) / h
    m = grid_view(rp, core_mask(rp, 2).astype(np.int8)).astype(bool)
    assert np.allclose(grid_view(rp, out[0])[m], expect[m], rtol=1e-12, atol=1e-16)


def test_energy_hyperdiffusion_acts_on_internal_energy():
    """hyperdifesource1 (src/hyperdifesource1.cu:18-554) diffuses
    q = e - (b^2 + m^2/rho_t)/2: adding uniform kinetic and magnetic energy
    that leaves q unchanged leaves the e-update unchanged up to the factor
    cmax of nu (which sees the new wave speeds).  q varies along dim 0 only,
    so only dim 0 contributes."""
    dt = 1.0e-3
    rp = box_params(ndim=2, ni=(16, 12), bc=BC_PERIODIC, terms=T_HE, dt=dt, chyp=0.05)
    rng = np.random.default_rng(9)
    w = uniform_background(rp, rho_b=1.0, p_b=0.6)
    line = 0.05 * rng.standard_normal(rp.n[0])
    w[3] = np.tile(line, rp.n[1])
    X = mesh(rp)
    w2 = w.copy()
    w2[1] = 0.25
    w2[4] = 0.5
    w2[3] = w[3] + 0.5 * (0.25 ** 2 + 0.5 ** 2)
    deltas, cms = [], []
    for st in (w, w2):
        o = Oracle(rp, st, X)
        cms.append(o.visc(3, 0)[2])
        o.close()
        deltas.append(one_step(rp, st, X)[3] - st[3])
    m = core_mask(rp, 2)
    assert np.abs(deltas[0][m]).max() > 0 and cms[1] > cms[0]
    assert np.allclose(deltas[1][m] * (cms[0] / cms[1]), deltas[0][m], rtol=1e-9, atol=1e-15)


def _line_state(rp, b_b=(0.0, 0.0, 0.0)):
    X = mesh(rp)
    w = uniform_background(rp, rho_b=1.0, p_b=0.6, b_b=b_b)
    return X, w


def test_total_pressure_in_momentum_flux():
    """Momentum flux along x with m = 0, b_x = b_bx = 0 is the total
    pressure p_t = (gamma-1) e - (gamma-2)(b.b_b + b^2/2) (computept3,
    src/dervfields.cu:107-153; computeflux, src/centdiff1.cu:51-219): for
    cubic p_t one Euler flux step gives dm_x = -dt dp_t/dx exactly, and no
    other field moves."""
    dt, B0, gam = 1.0e-3, 0.3, 5.0 / 3.0
    rp = box_params(ndim=2, ni=(12, 12), bc=BC_CONTCD4, terms=T_FLUX, dt=dt, hyper=0)
    X, w = _line_state(rp, b_b=(0.0, B0, 0.0))
    x = X[0]
    e = 1e-3 * (0.2 + x - 1.5 * x ** 2 + 2.0 * x ** 3)
    de = 1e-3 * (1.0 - 3.0 * x + 6.0 * x ** 2)
    by = 1e-3 * (0.5 - x)
    dby = -1e-3
    w[3], w[5] = e, by
    out = one_step(rp, w, X)
    dpt = (gam - 1) * de - (gam - 2) * (B0 * dby + by * dby)
    m = core_mask(rp, 4)
    assert np.allclose(out[1][m], -dt * dpt[m], rtol=1e-9, atol=1e-20)
    for f in (0, 2, 3, 4, 5):
        assert np.allclose(out[f][m], w[f][m], rtol=1e-15, atol=1e-20), f


def test_energy_flux_of_uniform_flow():
    """Uniform v0 = m/rho_t along x, b = 0: the energy flux is
    v0 (e + p_t + e_b) (computefluxe, src/centdiff2.cu:158-220), so
    de = -dt v0 gamma de/dx and dm_x = -dt (gamma-1) de/dx for cubic e."""
    dt, v0, gam = 1.0e-3, 0.2, 5.0 / 3.0
    rp = box_params(ndim=2, ni=(12, 12), bc=BC_CONTCD4, terms=T_FLUX, dt=dt, hyper=0)
    X, w = _line_state(rp)
    x = X[0]
    e = 1e-2 * (0.2 + x - 1.5 * x ** 2 + 2.0 * x ** 3)
    de = 1e-2 * (1.0 - 3.0 * x + 6.0 * x ** 2)
    w[1], w[3] = v0, e
    out = one_step(rp, w, X)
    m = core_mask(rp, 4)
    assert np.allclose(out[3][m] - w[3][m], -dt * v0 * gam * de[m], rtol=1e-8, atol=1e-20)
    assert np.allclose(out[1][m] - w[1][m], -dt * (gam - 1) * de[m], rtol=1e-8, atol=1e-20)
    assert np.allclose(out[0][m], 0.0, atol=1e-18)


def test_background_energy_terms():
    """addenergyterms (src/centdiff2.cu:80-115): with v = v_x(x) only,
    de = dt dv_x/dx (b_bx^2 - p_tb), p_tb = (gamma-1) e_b - (gamma-2) b_b^2/2."""
    dt, gam = 1.0e-3, 5.0 / 3.0
    bb = (0.4, 0.25, 0.0)
    rp = box_params(ndim=2, ni=(12, 12), bc=BC_CONTCD4, terms=T_ETERM, dt=dt, hyper=0)
    X, w = _line_state(rp, b_b=bb)
    x = X[0]
    w[1] = 1e-2 * (0.3 - x + 2.0 * x ** 2 - x ** 3)  # rho_t = 1: v_x = m_x
    dv = 1e-2 * (-1.0 + 4.0 * x - 3.0 * x ** 2)
    ptb = (gam - 1) * w[6] - 0.5 * (gam - 2) * (bb[0] ** 2 + bb[1] ** 2)
    out = one_step(rp, w, X)
    m = core_mask(rp, 4)
    assert np.allclose(out[3][m] - w[3][m], dt * dv[m] * (bb[0] ** 2 - ptb[m]), rtol=1e-9, atol=1e-20)


def test_fluxes_of_linear_flow():
    """rho_t = 1, uniform e = e0, b = 0, m = a + c x along x.  Fluxes
    (src/centdiff1.cu:51-219, src/centdiff2.cu:158-220): F_rho = m,
    F_mx = m^2 + p_t with p_t = (gamma-1)(e0 - m^2/2), F_e = m (e + p_t + e_b);
    all are polynomials of degree <= 3, so one Euler step is exact."""
    dt, gam, e0 = 1.0e-3, 5.0 / 3.0, 0.05
    rp = box_params(ndim=2, ni=(12, 12), bc=BC_CONTCD4, terms=T_FLUX, dt=dt, hyper=0)
    X, w = _line_state(rp)
    x = X[0]
    a, c = 0.02, -0.03
    mx = a + c * x
    w[1], w[3] = mx, e0
    eb = w[6]
    out = one_step(rp, w, X)
    m = core_mask(rp, 4)
    assert np.allclose(out[0][m], -dt * c, rtol=1e-9, atol=0)
    assert np.allclose(out[1][m] - w[1][m], -dt * (3.0 - gam) * mx[m] * c, rtol=1e-8, atol=1e-20)
    de = -dt * (c * (gam * e0 + eb[m]) - 1.5 * (gam - 1) * mx[m] ** 2 * c)
    assert np.allclose(out[3][m] - w[3][m], de, rtol=1e-8, atol=1e-20)


# ---------------------------------------------------------------- momentum and field hyperdiffusion

def _hyper_case(terms, field, g_line, h_col, rho_b=1.0):
    """Periodic 2D box (ghost rows hold images), rho_t = rho_b uniform,
    field = g(x) h(y); every other evolved field zero except e (pressure)."""
    rp = box_params(ndim=2, ni=(16, 12), bc=BC_PERIODIC, terms=terms, dt=1.0e-3, chyp=0.05)
    X, w = _line_state(rp, b_b=(0.2, 0.1, 0.0))
    w[9 - 2 * 0] = w[7]  # keep b_b as set (index 8,9 are b_b in 2D)
    w[7] = rho_b
    gv = grid_view(rp, w)
    gv[field] = np.outer(h_col, g_line)
    return rp, X, w


def _nu(rp, w, X, field, dim):
    o = Oracle(rp, w, X)
    st = o.state()
    nur, nul, _ = o.visc(field, dim)
    o.close()
    return st, grid_view(rp, nur), grid_view(rp, nul)


def _smooth_line(n, seed, band=6):
    """Random smooth profile that vanishes within `band` cells of both ends."""
    rng = np.random.default_rng(seed)
    g = np.convolve(rng.standard_normal(n + 6), np.ones(4) / 4, mode="same")[3:-3]
    g[:band] = 0.0
    g[n - band:] = 0.0
    return 0.01 * g


def test_momentum_hyperdiffusion_diagonal_closed_form():
    """hyperdifmomsource1 (src/hyperdifmomsource1.cu:17-535), diagonal term:
    with rho_t = 1 and v_x = v(x) only, dm_x = dt (nu_r D+v - nu_l D-v)/dx^2
    (a conservative discretisation of d/dx(rho nu dv/dx)) and
    de = dt (m_r nu_r D+v - m_l nu_l D-v)/dx^2 with face-averaged m."""
    from smaug_amd.params import T_HMOM
    rp = box_params(ndim=2, ni=(16, 12), bc=BC_PERIODIC, terms=T_HMOM, dt=1.0e-3, chyp=0.05)
    X, w = _line_state(rp, b_b=(0.2, 0.1, 0.0))
    w[0] = np.tile(10.0 * _smooth_line(rp.n[0], 20, band=0), rp.n[1])  # rho_t = 1 + rho'(x)
    w[1] = (1.0 + w[0]) * np.tile(_smooth_line(rp.n[0], 21, band=0), rp.n[1])
    st, nur, nul = _nu(rp, w, X, 1, 0)
    out = grid_view(rp, one_step(rp, w, X))
    rt = 1.0 + grid_view(rp, st[0])
    mx = grid_view(rp, st[1])
    v = mx / rt
    h = rp.dx[0]
    sh = lambda a, s: np.roll(a, -s, axis=1)  # noqa: E731  a[i+s] along x
    rr, rl = 0.5 * (rt + sh(rt, 1)), 0.5 * (rt + sh(rt, -1))
    mr, ml = 0.5 * (mx + sh(mx, 1)), 0.5 * (mx + sh(mx, -1))
    gr, gl = sh(v, 1) - v, v - sh(v, -1)
    dm = rp.dt * (rr * nur * gr - rl * nul * gl) / h ** 2
    de = rp.dt * (mr * nur * gr - ml * nul * gl) / h ** 2
    m = grid_view(rp, core_mask(rp, 2).astype(np.int8)).astype(bool)
    assert np.abs(dm[m]).max() > 0
    assert np.allclose((out[1] - mx)[m], dm[m], rtol=1e-9, atol=1e-22)
    assert np.allclose((out[3] - grid_view(rp, st[3]))[m], de[m], rtol=1e-9, atol=1e-22)
    # the diagonal term dissipates kinetic energy: sum v dm = -sum rho nu (Dv)^2 < 0
    assert (v[m] * dm[m]).sum() < 0


def test_momentum_hyperdiffusion_off_diagonal_closed_form():
    """hyperdifmomsourcene1 (src/hyperdifmomsourcene1.cu:17-752): v_y =
    g(x) h(y) with h linear.  nu of v_y along x is then the same on every
    row (a ratio of differences), and the cross term adds
    dm_x = dt d/dy[rho_t (nu_r + nu_l)/4 dv_y/dx] (central differences),
    while v_y gets half the diagonal form along x."""
    from smaug_amd.params import T_HMOM
    rp = box_params(ndim=2, ni=(16, 12), bc=BC_PERIODIC, terms=T_HMOM, dt=1.0e-3, chyp=0.05)
    X, w = _line_state(rp, b_b=(0.2, 0.1, 0.0))
    g = _smooth_line(rp.n[0], 22, band=0) + 0.02
    hcol = 1.0 + 0.5 * np.arange(rp.n[1]) / rp.n[1]
    w[0] = np.tile(10.0 * _smooth_line(rp.n[0], 24, band=0), rp.n[1])  # rho_t = 1 + rho'(x)
    gv = grid_view(rp, w)
    gv[2] = (1.0 + gv[0]) * np.outer(hcol, g)
    st, nur, nul = _nu(rp, w, X, 2, 0)
    _, nury, nuly = _nu(rp, w, X, 2, 1)  # along y: nonzero only where ghost images break linearity
    out = grid_view(rp, one_step(rp, w, X))
    rt = 1.0 + grid_view(rp, st[0])
    my = grid_view(rp, st[2])
    v = my / rt
    hx, hy = rp.dx
    sx = lambda a, s: np.roll(a, -s, axis=1)  # noqa: E731
    sy = lambda a, s: np.roll(a, -s, axis=0)  # noqa: E731
    t6 = (sx(v, 1) - sx(v, -1)) / (2 * hx) * (nur + nul) / 4.0
    ga = rt * t6
    dmx = rp.dt * (sy(ga, 1) - sy(ga, -1)) / (2 * hy)
    rr, rl = 0.5 * (rt + sx(rt, 1)), 0.5 * (rt + sx(rt, -1))
    dmy = rp.dt * (rr * nur * (sx(v, 1) - v) - rl * nul * (v - sx(v, -1))) / (2 * hx ** 2)
    ry, ly = 0.5 * (rt + sy(rt, 1)), 0.5 * (rt + sy(rt, -1))
    dmy += rp.dt * (ry * nury * (sy(v, 1) - v) - ly * nuly * (v - sy(v, -1))) / hy ** 2
    m = grid_view(rp, core_mask(rp, 3).astype(np.int8)).astype(bool)
    assert np.abs(dmx[m]).max() > 0
    assert np.allclose(out[1][m], dmx[m], rtol=1e-8, atol=1e-20)
    assert np.allclose((out[2] - my)[m], dmy[m], rtol=1e-8, atol=1e-20)


def test_field_hyperdiffusion_closed_form():
    """hyperdifbsource1 / hyperdifbsourcene1 (src/hyperdifbsource1.cu:17-448,
    src/hyperdifbsourcene1.cu:17-829): b_y = g(x) h(y), h linear.  Diagonal:
    db_y = dt (nu_r D+b_y - nu_l D-b_y)/dx^2; off-diagonal (curl form):
    db_x = -dt d/dy[(nu_r + nu_l)/2 db_y/dx] with central differences."""
    from smaug_amd.params import T_HB
    rp = box_params(ndim=2, ni=(16, 12), bc=BC_PERIODIC, terms=T_HB, dt=1.0e-3, chyp=0.05)
    X, w = _line_state(rp, b_b=(0.2, 0.1, 0.0))
    g = _smooth_line(rp.n[0], 23, band=0) + 0.02
    hcol = 1.0 + 0.5 * np.arange(rp.n[1]) / rp.n[1]
    gv = grid_view(rp, w)
    gv[5] = np.outer(hcol, g)
    st, nur, nul = _nu(rp, w, X, 5, 0)
    out = grid_view(rp, one_step(rp, w, X))
    b = grid_view(rp, st[5])
    hx, hy = rp.dx
    bp, bm = np.roll(b, -1, axis=1), np.roll(b, 1, axis=1)
    t3 = (bp - bm) / (2 * hx) * (nur + nul) / 2.0
    dbx = -rp.dt * (np.roll(t3, -1, axis=0) - np.roll(t3, 1, axis=0)) / (2 * hy)
    dby = rp.dt * (nur * (bp - b) - nul * (b - bm)) / hx ** 2
    m = grid_view(rp, core_mask(rp, 3).astype(np.int8)).astype(bool)
    assert np.abs(dbx[m]).max() > 0
    assert np.allclose(out[4][m], dbx[m], rtol=1e-8, atol=1e-20)
    assert np.allclose((out[5] - b)[m], dby[m], rtol=1e-8, atol=1e-20)


@pytest.mark.parametrize("terms_name", ["hmom", "hb"])
def test_momentum_and_field_hyperdiffusion_conserve(terms_name):
    """With the perturbation supported in the core (zero within 6 cells of
    every face, so no halo rule is involved), the m and b hyperdiffusion
    sources are in flux form: the interior sums of m and b are unchanged."""
    from smaug_amd.params import T_HB, T_HMOM
    T = T_HMOM if terms_name == "hmom" else T_HB
    rp = box_params(ndim=2, ni=(20, 20), bc=BC_PERIODIC, terms=T, dt=1.0e-3, chyp=0.05)
    X, w = _line_state(rp, b_b=(0.2, 0.1, 0.0))
    gv = grid_view(rp, w)
    for f, seed in ((1, 31), (2, 32), (4, 33), (5, 34)):
        gv[f] = np.outer(_smooth_line(rp.n[1], seed), _smooth_line(rp.n[0], seed + 10))
    out = one_step(rp, w, X)
    m = core_mask(rp, 2)
    moved = 0.0
    for f in (1, 2, 4, 5):
        moved = max(moved, np.abs(out[f][m] - w[f][m]).max())
        assert abs(out[f][m].sum() - w[f][m].sum()) <= 1e-14 * np.abs(w[f][m]).sum(), f
    assert moved > 0

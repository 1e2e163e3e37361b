"""Pins of the oracle's reference quirks at the domain faces (no GPU).

The closed-form pins (tests/test_oracle_pins.py) check cells away from the
faces.  The cells next to a non-periodic face, and the hyperdiffusion
windows that reach past the grid, follow rules of the reference that no
closed form fixes; here each rule is evaluated independently, by brute
force from the reference text, on every cell of small grids and compared
with the oracle:

  Q1   the zeroing masks of grad3dn / grad1l3n / grad1r3n / grad13n with
       C precedence (A || B && C), src/gradops.cu:437-823;
  Q2   the flux divergence reads boundtype[] at the DEV index of `flux`,
       src/centdiff1.cu:26, src/gradops.cu:463;
  Q3   v is formed on cells transverse-interior to the sweep only, so the
       energy terms' cross-gradients read zeros off it (periodic grids),
       src/computedervfields.cu:82-105, src/centdiff2.cu:80-115;
  Q7   the right-chain halo of hyperdifvisc1r: a per-thread emulation, in
       increasing thread index, of the zeroing and the stray write with the
       unpadded encoder, src/hyperdifvisc1r.cu:1099-1253;
  Q8   the left-chain halo padded 0 <- 6, src/hyperdifvisc1l.cu:17-60;
  Q11  extrapolated ghost positions and delx on a non-uniform grid,
       src/init.cu:2622-3010;
  Q15  ghost-cell increments of the hyperdiffusion sources never survive
       the boundary pass (ghost cells equal their images after a step);
and the hydrostatic background of the stratified configurations against
an independent quadrature of dp/dz = -rho g.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle.pyoracle import Oracle
from smaug_amd import configs
from smaug_amd.params import (BC_CONTCD4, BC_PERIODIC, T_ETERM, T_FLUX, RunParams, npert, nvar)

BC_CONT, BC_FIXED, BC_SYMM, BC_ASYMM = 3, 5, 6, 7


def box(ndim, ni, bc, terms=0xFF, dt=1e-3, chyp=0.05, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    return RunParams(ndim=ndim, ni=list(ni[:ndim]), xmin=list(lo[:ndim]), xmax=list(hi[:ndim]), dt=dt,
                     gamma=5.0 / 3.0, g=[0.0, 0.0, 0.0], courant=0.3, rkon=0, moddton=0, hyperdifmom=1,
                     chyp={"default": chyp}, bcu=[bc] * 3, bcl=[bc] * 3, terms=terms)


def background(rp, rho_b=1.0, p_b=0.6, b_b=(0.0, 0.0, 0.0)):
    nd = rp.ndim
    w = np.zeros((nvar(nd), rp.ncells))
    w[2 + 2 * nd] = p_b / (rp.gamma - 1.0) + 0.5 * sum(b * b for b in b_b[:nd])
    w[3 + 2 * nd] = rho_b
    for d in range(nd):
        w[4 + 2 * nd + d] = b_b[d]
    return w


def coords(rp):
    """Integer cell coordinates (ndim, N), i fastest."""
    n = rp.n
    idx = np.arange(rp.ncells)
    out = [idx % n[0], (idx // n[0]) % n[1]]
    if rp.ndim == 3:
        out.append(idx // (n[0] * n[1]))
    return np.stack(out)


# ---------------------------------------------------------------- Q1

def _inside(x, lo, hi):
    """lo < x < hi, the reference's strict bounds."""
    return (x > lo) & (x < hi)


def grad3dn_mask(rp, d, nonper):
    """(computed, zeroed) per cell for grad3dn along d (src/gradops.cu:437-602):
    computed for 1 < x < n-2; when boundtype != 0 zeroed at n-3 always, at
    n-4 when every transverse index t has 1 < t < n-2 (the C precedence
    `x==n-3 || x==n-4 && trans`), and likewise at 2 / 3."""
    X = coords(rp)
    n = rp.n
    x = X[d]
    computed = _inside(x, 1, n[d] - 2)
    trans = np.ones(rp.ncells, dtype=bool)
    for t in range(rp.ndim):
        if t != d:
            trans &= _inside(X[t], 1, n[t] - 2)
    zeroed = np.zeros(rp.ncells, dtype=bool)
    if nonper:
        zeroed = (x == n[d] - 3) | ((x == n[d] - 4) & trans) | (x == 2) | ((x == 3) & trans)
    return computed, zeroed


@pytest.mark.parametrize("ndim,bc", [(2, BC_CONTCD4), (2, BC_PERIODIC), (3, BC_CONTCD4)])
def test_q1_grad3dn_zeroing_on_every_cell(ndim, bc):
    """A quartic along d times a cubic across it: grad3dn is its exact
    derivative wherever it is computed and not zeroed, and exactly 0 where
    the reference's mask (with its operator precedence) zeroes it or where
    it is not computed."""
    rp = box(ndim, (10, 9, 8), bc)
    pos = np.stack(configs._mesh(rp))
    o = Oracle(rp, background(rp), pos)
    for d in range(ndim):
        x = pos[d]
        t = pos[(d + 1) % ndim]
        a = (0.3 - 1.1 * x + 0.7 * x ** 2 - 2.0 * x ** 3 + 1.5 * x ** 4) * (1.0 + t + t ** 3)
        da = (-1.1 + 1.4 * x - 6.0 * x ** 2 + 6.0 * x ** 3) * (1.0 + t + t ** 3)
        g = o.grad("grad3dn", a, 0, d)
        computed, zeroed = grad3dn_mask(rp, d, bc != BC_PERIODIC)
        live = computed & ~zeroed
        assert np.all(g[~live] == 0.0), d
        assert np.allclose(g[live], da[live], rtol=1e-10, atol=1e-10), d
        assert live.sum() > 0 and (computed & zeroed).sum() > 0 or bc == BC_PERIODIC
    o.close()


def grad1_mask(rp, kind, d, nonper):
    """(computed, zeroed) for grad1l3n / grad1r3n / grad13n along d
    (src/gradops.cu:608-823): computed for x > 0 (1l), x < n-1 (1r),
    0 < x < n-1 (13n); when boundtype != 0 zeroed at n-2 always, at n-3
    when the transverse indices are inside (0, n-1), and at 1 / 2 likewise,
    except two asymmetric transverse bounds in 3D: k > 1 at n-3 for
    grad1l3n along dim 0 (:627), j > 1 at 2 for grad13n along dim 2 (:816)."""
    X = coords(rp)
    n = rp.n
    x = X[d]
    if kind == "grad1l3n":
        computed = x > 0
    elif kind == "grad1r3n":
        computed = x < n[d] - 1
    else:
        computed = (x > 0) & (x < n[d] - 1)

    def trans(side):
        m = np.ones(rp.ncells, dtype=bool)
        for t in range(rp.ndim):
            if t == d:
                continue
            lo = 0
            if rp.ndim == 3 and kind == "grad1l3n" and d == 0 and side == "hi" and t == 2:
                lo = 1
            if rp.ndim == 3 and kind == "grad13n" and d == 2 and side == "lo" and t == 1:
                lo = 1
            m &= _inside(X[t], lo, n[t] - 1)
        return m

    zeroed = np.zeros(rp.ncells, dtype=bool)
    if nonper:
        zeroed = ((x == n[d] - 2) | ((x == n[d] - 3) & trans("hi")) | (x == 1) | ((x == 2) & trans("lo")))
    return computed, zeroed


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("kind", ["grad1l3n", "grad1r3n", "grad13n"])
def test_q1_first_order_zeroing_on_every_cell(ndim, kind):
    """Quadratic along d times a function across it: the one-sided and
    central first differences are exact at x -+ h/2 and x, and exactly 0
    where the reference's masks zero them, including the two asymmetric
    transverse bounds of the 3D operators."""
    rp = box(ndim, (10, 9, 8), BC_CONTCD4)
    pos = np.stack(configs._mesh(rp))
    o = Oracle(rp, background(rp), pos)
    for d in range(ndim):
        x = pos[d]
        t = pos[(d + 1) % ndim] + 2.0 * pos[(d + 2) % ndim] if ndim == 3 else pos[(d + 1) % ndim]
        h = rp.dx[d]
        s = 1.0 + t * t
        a = (2.0 * x ** 2 - 3.0 * x + 0.25) * s
        fp = lambda z: (4.0 * z - 3.0) * s  # noqa: E731
        want = {"grad1l3n": fp(x - h / 2), "grad1r3n": fp(x + h / 2), "grad13n": fp(x)}[kind]
        g = o.grad(kind, a, 0, d)
        computed, zeroed = grad1_mask(rp, kind, d, True)
        live = computed & ~zeroed
        assert np.all(g[~live] == 0.0), (kind, d)
        assert np.allclose(g[live], want[live], rtol=1e-11, atol=1e-11), (kind, d)
    o.close()
    if ndim == 3:  # the asymmetric bounds make a difference on these grids
        c1, z1 = grad1_mask(rp, "grad1l3n", 0, True)
        c2, z2 = grad1_mask(rp, "grad1r3n", 0, True)
        assert np.any(z1 != z2)


# ---------------------------------------------------------------- Q2

def test_q2_flux_divergence_reads_the_flux_row_of_boundtype():
    """2D, rho_t = 1, momentum m_0 cubic in x, b_2 uniform, fixed boundaries
    (the ghosts keep the polynomial) on every field except the BC row at the
    DEV index of `flux` (= 2, the mom2 row), set periodic along dim 0: the
    mass update is -dt dm_0/dx and the b_2 update -dt b_2 dm_0/dx on every
    interior cell, unzeroed (src/centdiff1.cu:26 and src/centdiff2.cu:62
    pass `flux` as the field of grad3dn).  With that row fixed as well the
    faces' layers are zeroed."""
    for row_periodic in (True, False):
        rp = box(2, (12, 10), BC_FIXED, terms=T_FLUX, dt=1e-3)
        rp.bcu = [BC_FIXED, BC_PERIODIC, BC_PERIODIC]
        rp.bcl = [BC_FIXED, BC_PERIODIC, BC_PERIODIC]
        if row_periodic:  # the mom2 row: field index 2 = the DEV value of `flux` in 2D
            rp.bc_field = {"mom2": [BC_PERIODIC, BC_PERIODIC, BC_PERIODIC]}
        pos = np.stack(configs._mesh(rp))
        w = background(rp, rho_b=1.0, p_b=0.6)
        x = pos[0]
        m0 = 0.2 + 0.3 * x - 0.5 * x ** 2 + 0.4 * x ** 3
        w[1] = m0
        w[5] = 0.7  # b_2 uniform: its flux along dim 0 is b_2 v_0 = 0.7 m_0 (centdiff2)
        o = Oracle(rp, w, pos)
        o.step(1)
        st = o.state()
        drho = (st[0] - w[0]) / rp.dt
        db2 = (st[5] - w[5]) / rp.dt
        o.close()
        X = coords(rp)
        n = rp.n
        interior = (X[0] >= 2) & (X[0] <= n[0] - 3) & (X[1] >= 2) & (X[1] <= n[1] - 3)
        exact = -(0.3 - 1.0 * x + 1.2 * x ** 2)
        computed, zeroed = grad3dn_mask(rp, 0, not row_periodic)
        live = interior & ~zeroed
        assert np.allclose(drho[live], exact[live], rtol=1e-9, atol=1e-9)
        assert np.all(np.abs(drho[interior & zeroed]) < 1e-12)
        assert np.allclose(db2[live], 0.7 * exact[live], rtol=1e-9, atol=1e-9)
        assert np.all(np.abs(db2[interior & zeroed]) < 1e-12)
        if row_periodic:
            assert not np.any(interior & zeroed)
        else:
            assert np.any(interior & zeroed)



# ---------------------------------------------------------------- Q3

def test_q3_velocity_masks_leave_the_interior_updates_alone():
    """computevels fills v only on cells transverse-interior to its sweep
    (src/computedervfields.cu:82-105), but the loop body runs it for every
    sweep first (src/iosmaug.c:637-640), so v is 0 only on cells with two
    ghost coordinates, which no stencil of an interior cell reads.  2D
    periodic, uniform along y, background field with b_bx b_by != 0: the
    energy-term cross-gradient d v_0/dy (addenergyterms, src/centdiff2.cu:
    80-115) reads the ghost rows j = 0, 1, n1-2, n1-1 of interior columns.
    With those rows filled (the union of the sweeps) it is exactly 0; a
    single sweep's mask would leave (7, -1, 1, -7) v/(12 dy) on rows 2, 3,
    n1-4, n1-3.  With v = v(x) every interior row then changes alike."""
    rp = box(2, (12, 10), BC_PERIODIC, terms=T_ETERM, dt=1e-3)
    pos = np.stack(configs._mesh(rp))
    bbx, bby = 0.3, 0.4
    X = coords(rp)
    n0, n1 = rp.n[0], rp.n[1]
    interior = (X[0] >= 2) & (X[0] <= n0 - 3) & (X[1] >= 2) & (X[1] <= n1 - 3)
    for vfun in (lambda x: 0.01 + 0.0 * x, lambda x: 0.01 * (1.0 + 0.5 * np.sin(2 * np.pi * x))):
        w = background(rp, rho_b=1.0, p_b=0.6, b_b=(bbx, bby, 0.0))
        w[1] = vfun(pos[0])  # rho_t = 1, so m_0 = v_0
        o = Oracle(rp, w, pos)
        o.step(1)
        de = (o.state()[3] - w[3]) / rp.dt
        o.close()
        rows = [np.where(interior & (X[1] == j))[0] for j in range(2, n1 - 2)]
        for r in rows[1:]:
            assert np.array_equal(de[r], de[rows[0]])
    # uniform flow: no gradient in any direction, so no energy term at all
    assert np.all(de[interior] != 0.0)  # (the last case has an along-x term)
    w = background(rp, rho_b=1.0, p_b=0.6, b_b=(bbx, bby, 0.0))
    w[1] = 0.01
    o = Oracle(rp, w, pos)
    o.step(1)
    de0 = (o.state()[3] - w[3]) / rp.dt
    o.close()
    assert np.all(de0[interior] == 0.0)


# ---------------------------------------------------------------- Q7 / Q8

def right_halo(rp, q):
    """Per-thread emulation of hyperdifvisc1r_parallel (src/hyperdifvisc1r.cu:
    1099-1253), threads in increasing linear index, each running its whole
    body: phase 1 zeroes its padded cell (and the far halo layers when it
    is the last along a dim), phase 2 stores q with the unpadded encoder
    encode3(i+1, j+1[, k+1]) into the padded array.  hyperdifvisc1ar then
    overwrites the padded interior with q.  Returns the padded array."""
    nd = rp.ndim
    n = list(rp.n[:nd])
    P = [m + 2 for m in n]
    pad = np.full(int(np.prod(P)), np.nan)  # whatever an earlier launch left

    def e2(i, j, k=0):
        return (k * P[1] + j) * P[0] + i

    def e3(i, j, k=0):
        return (k * n[1] + j) * n[0] + i

    qq = q.reshape(tuple(n[::-1]))
    n2 = n[2] if nd == 3 else 1
    for k in range(n2):
        for j in range(n[1]):
            for i in range(n[0]):
                li, lj, lk = i == n[0] - 1, j == n[1] - 1, nd == 3 and k == n2 - 1
                zero = [(i, j, k)]
                if li:
                    zero += [(i + 1, j, k), (i + 2, j, k)]
                if lj:
                    zero += [(i, j + 1, k), (i, j + 2, k)]
                if lk:
                    zero += [(i, j, k + 1), (i, j, k + 2)]
                if li and lj:
                    zero += [(i + 1 + a, j + 1 + b, k) for a in (0, 1) for b in (0, 1)]
                if li and lk:
                    zero += [(i + 1 + a, j, k + 1 + c) for a in (0, 1) for c in (0, 1)]
                if lj and lk:
                    zero += [(i, j + 1 + b, k + 1 + c) for b in (0, 1) for c in (0, 1)]
                if li and lj and lk:
                    zero += [(i + 1 + a, j + 1 + b, k + 1 + c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]
                for z in zero:
                    pad[e2(*z)] = 0.0
                val = qq[k, j, i] if nd == 3 else qq[j, i]
                pad[e3(i + 1, j + 1, k + 1 if nd == 3 else 0)] = val
    for k in range(n2):
        for j in range(n[1]):
            for i in range(n[0]):
                pad[e2(i + 1, j + 1, k + 1 if nd == 3 else 0)] = qq[k, j, i] if nd == 3 else qq[j, i]
    return pad.reshape(tuple(P[::-1]))


def nu_windows(line, dxk):
    """nu_{i+1/2}, nu_{i-1/2} of cell i from the extended line q(-1..n)
    (index 0 of `line` is q(-1)): max |3rd difference| over 3 interfaces /
    max |1st difference| over 5 (src/hyperdifvisc1r.cu:426-843)."""
    def q(i):
        return line[i + 1]

    def d3(f):  # third difference across interface f+1/2
        return abs(3.0 * (q(f + 1) - q(f)) - (q(f + 2) - q(f - 1)))

    def d1(f):
        return abs(q(f + 1) - q(f))

    n = len(line) - 2
    nur, nul = np.zeros(n), np.zeros(n)
    for i in range(2, n - 2):
        for face, out in ((i, nur), (i - 1, nul)):
            num = max(d3(face + m) for m in (-1, 0, 1))
            den = max(d1(face + m) for m in (-2, -1, 0, 1, 2))
            out[i] = dxk[i] * num / den if den > 0 else 0.0
    return nur, nul


@pytest.mark.parametrize("ndim,bc", [(2, BC_CONTCD4), (2, BC_PERIODIC), (3, BC_CONTCD4)])
@pytest.mark.parametrize("field", [0, 1])
def test_q7_q8_nu_on_every_interior_cell(ndim, bc, field):
    """nu_r / nu_l of rho and of v_0 = m_0/rho_t at every interior cell,
    faces included, against windows on the line extended by the halo values
    of the two chains: q(n) = the right chain's padded halo from the
    per-thread emulation of the race (Q7; 0 or a stray q), q(-1) = q(5), the
    left chain's halo fix (Q8).  The right chain never reads q(-1) and the
    left chain never reads q(n)."""
    rng = np.random.default_rng(5 + field + 10 * ndim)
    ni = (10, 9, 8) if ndim == 3 else (13, 10)
    rp = box(ndim, ni, bc, chyp=0.05)
    pos = np.stack(configs._mesh(rp))
    w = background(rp, rho_b=1.0, p_b=0.6, b_b=(0.2, 0.1, 0.05))
    w[0] = 0.05 * rng.standard_normal(rp.ncells)
    w[1] = 0.03 * rng.standard_normal(rp.ncells)
    o = Oracle(rp, w, pos)
    st = o.state()  # ghosts filled by the boundary pass
    _, delx = o.grid()
    q = st[0] if field == 0 else st[1] / (st[0] + st[3 + 2 * ndim])
    pad = right_halo(rp, q)
    shape = tuple(rp.n[:ndim][::-1])
    Q = q.reshape(shape)
    for d in range(ndim):
        nur, nul, cmax = o.visc(field, d)
        NR, NL = nur.reshape(shape), nul.reshape(shape)
        DX = delx[d].reshape(shape)
        ax = ndim - 1 - d
        n = rp.n[d]
        checked = 0
        for idx in np.ndindex(*[s for a, s in enumerate(shape) if a != ax]):
            full = list(idx)
            full.insert(ax, slice(None))
            tr = [t for a, t in enumerate(full) if a != ax]
            dims_t = [dd for dd in range(ndim - 1, -1, -1) if dd != d]
            if not all(2 <= tt <= rp.n[dd] - 3 for tt, dd in zip(tr, dims_t)):
                continue
            line = Q[tuple(full)]
            ppos = [t + 1 for t in full if not isinstance(t, slice)]
            pfull = list(ppos)
            pfull.insert(ax, n + 1)
            halo_r = pad[tuple(pfull)]
            ext_r = np.concatenate([[np.nan], line, [halo_r]])
            ext_l = np.concatenate([[line[5]], line, [np.nan]])
            dxk = DX[tuple(full)] * cmax * 0.05
            br, _ = nu_windows(ext_r, dxk)
            _, bl = nu_windows(ext_l, dxk)
            sl = slice(2, n - 2)
            assert np.allclose(NR[tuple(full)][sl], br[sl], rtol=1e-12, atol=0), (d, idx)
            assert np.allclose(NL[tuple(full)][sl], bl[sl], rtol=1e-12, atol=0), (d, idx)
            checked += 1
        assert checked > 0
    o.close()


def test_q7_stray_halo_is_not_all_zero():
    """The emulated race leaves stray q values in the right chain's padded
    halo on these grids (otherwise the test above would not exercise Q7),
    and zeros elsewhere in it."""
    rp = box(2, (13, 10), BC_CONTCD4)
    q = np.arange(1.0, rp.ncells + 1.0)
    pad = right_halo(rp, q)
    halo = pad[1:-1, -1]
    assert np.any(halo != 0.0) and np.any(halo == 0.0)
    assert np.all(np.isfinite(pad))


# ---------------------------------------------------------------- Q11

@pytest.mark.parametrize("ndim", [2, 3])
def test_q11_ghost_positions_are_linear_extrapolations(ndim):
    """initgrid (src/init.cu:2622-3010) on a quadratic, non-uniform grid
    whose ghost positions are garbage: upper ghosts and the padded layer
    past them are the linear extrapolations 2x(n-3)-x(n-4), 3x(n-3)-2x(n-4),
    4x(n-3)-3x(n-4); lower ones in dims 1 and 2 likewise from x(2), x(3)
    over three layers, but in dim 0 over two layers from x(1), x(2) (ghost
    cell 1 keeps its value); delx = (x(i+1) - x(i-1))/2 with those values."""
    ni = (10, 9, 8)[:ndim]
    rp = box(ndim, ni, BC_CONTCD4)
    n = rp.n
    axes = []
    rng = np.random.default_rng(3)
    for d in range(ndim):
        i = np.arange(n[d], dtype=np.float64)
        x = 0.1 + 0.5 * i + 0.03 * i * i + 0.01 * d
        x[[0, 1, n[d] - 2, n[d] - 1]] = rng.uniform(-9.0, 9.0, 4)  # garbage ghosts
        axes.append(x)
    grids = np.meshgrid(*axes[::-1], indexing="ij")[::-1]
    pos = np.stack([g.ravel() for g in grids])
    o = Oracle(rp, background(rp), pos)
    gpos, gdx = o.grid()
    o.close()
    for d in range(ndim):
        x = axes[d].copy()
        m = n[d]
        ext = np.empty(m + 2)  # index 0 is x(-1), index m+1 is x(m)
        ext[1:m + 1] = x
        ext[m - 1] = 2 * x[m - 3] - x[m - 4]      # x(n-2)
        ext[m] = 3 * x[m - 3] - 2 * x[m - 4]      # x(n-1)
        ext[m + 1] = 4 * x[m - 3] - 3 * x[m - 4]  # x(n)
        if d == 0:
            ext[1] = 2 * x[1] - x[2]              # x(0)
            ext[0] = 3 * x[1] - 2 * x[2]          # x(-1)
        else:
            ext[2] = 2 * x[2] - x[3]              # x(1)
            ext[1] = 3 * x[2] - 2 * x[3]          # x(0)
            ext[0] = 4 * x[2] - 3 * x[3]          # x(-1)
        want_pos = ext[1:m + 1]
        want_dx = 0.5 * (ext[2:m + 2] - ext[0:m])
        shape = tuple(n[:ndim][::-1])
        ax = ndim - 1 - d
        P = np.moveaxis(gpos[d].reshape(shape), ax, -1).reshape(-1, m)
        D = np.moveaxis(gdx[d].reshape(shape), ax, -1).reshape(-1, m)
        assert np.allclose(P, want_pos[None, :], rtol=1e-14, atol=1e-14), d
        assert np.allclose(D, want_dx[None, :], rtol=1e-13, atol=1e-14), d


# ---------------------------------------------------------------- Q15

IMAGE = {
    BC_CONTCD4: lambda x, n, nd: 4 - x if x < 2 else 2 * n - 6 - x,
    BC_CONT: lambda x, n, nd: 2 if x < 2 else n - 3,
    BC_SYMM: lambda x, n, nd: (3 - x if x < 2 else 2 * n - 5 - x) if nd == 2 else (4 - x if x < 2 else 2 * n - 6 - x),
}


@pytest.mark.parametrize("bc", [BC_CONTCD4, BC_CONT, BC_SYMM, BC_ASYMM])
def test_q15_ghost_increments_do_not_survive_the_boundary_pass(bc):
    """The hyperdiffusion sources add increments on ghost layers (they
    update every cell, src/hyperdifrhosource1.cu:83-95,204-215); after a
    full step every ghost cell of an image-type boundary equals its image
    (with the asymm sign), so none of them is visible."""
    rp, w, pos = configs.stratified_params(ni=(10, 8), ndim=2, rkon=1), None, None
    rp.bcu = [bc] * 3
    rp.bcl = [bc] * 3
    w, pos = configs.stratified(rp, perturb=11, amp=1e-3)
    o = Oracle(rp, w, pos)
    o.step(2)
    st = o.state()
    o.close()
    n0, n1 = rp.n[0], rp.n[1]
    sg = -1.0 if bc == BC_ASYMM else 1.0
    img = IMAGE[BC_SYMM if bc == BC_ASYMM else bc]
    for f in range(npert(2)):
        a = st[f].reshape(n1, n0)
        for j in range(2, n1 - 2):  # ghosts along dim 0 on interior rows
            for x in (0, 1, n0 - 2, n0 - 1):
                assert a[j, x] == sg * a[j, img(x, n0, 2)], (f, j, x)
        for i in range(n0):  # ghosts along dim 1 (the last pass) on every column
            for y in (0, 1, n1 - 2, n1 - 1):
                assert a[y, i] == sg * a[img(y, n1, 2), i], (f, i, y)


# ---------------------------------------------------------------- background

def test_hydrostatic_background_against_quadrature():
    """configs.hydrostatic (p = p0 exp(-mu g/R int dz/T), closed-form
    antiderivative of 1/T) against an independent adaptive quadrature of
    1/T(z); rho = p mu / (R T); rho(z0) = 2.7e-4; T(z) the tanh profile;
    and the centred difference of p balancing rho g to the difference's
    own truncation error (dp/dz = -rho g)."""
    from scipy.integrate import quad

    R, MU, G = 8.31e3, 1.2, 274.0
    z0 = 0.1e6
    z = np.linspace(0.1e6, 6.0e6, 61)
    p, rho = configs.hydrostatic(z)

    def T(zz):
        return 6000.0 + (1.0e6 - 6000.0) * (1.0 + math.tanh((zz - 2.0e6) / 0.2e6)) / 2.0

    p0 = 2.7e-4 * R * T(z0) / MU
    for zz, pp, rr in zip(z, p, rho):
        integral, _ = quad(lambda s: 1.0 / T(s), z0, zz, epsabs=0, epsrel=1e-13, limit=400, points=[2.0e6])
        want = p0 * math.exp(-(MU * G / R) * integral)
        assert pp == pytest.approx(want, rel=1e-10)
        assert rr == pytest.approx(pp * MU / (R * T(zz)), rel=1e-13)
    assert rho[0] == pytest.approx(2.7e-4, rel=1e-14)
    assert np.allclose(configs.temperature(z), [T(zz) for zz in z], rtol=1e-15)
    # dp/dz + rho g = 0: 4th-order centred difference on a fine grid
    h = 500.0
    zf = np.linspace(0.5e6, 5.5e6, 41)
    pp = [configs.hydrostatic(zf + s * h)[0] for s in (-2, -1, 1, 2)]
    dpdz = (pp[0] - 8 * pp[1] + 8 * pp[2] - pp[3]) / (12 * h)
    _, rf = configs.hydrostatic(zf)
    assert np.allclose(dpdz, -rf * G, rtol=1e-7)


def test_stratified_state_background_fields():
    """The stratified state's background: e_b = p/(gamma-1) + b_b^2/2 with
    b_b = 0.01 T / sqrt(mu0) along dim 0 only, rho_b = rho(z), zero
    perturbation when none is asked for."""
    rp = configs.stratified_params(ni=(12, 8, 8), ndim=3, rkon=1)
    w, pos = configs.stratified(rp)
    nd = 3
    z = pos[0]
    p, rho = configs.hydrostatic(z)
    bz = 0.01 / math.sqrt(4.0e-7 * math.pi)
    assert bz == pytest.approx(8.920620580763856, rel=1e-12)
    assert np.allclose(w[3 + 2 * nd], rho, rtol=1e-15)
    assert np.allclose(w[2 + 2 * nd], p / (rp.gamma - 1.0) + 0.5 * bz * bz, rtol=1e-15)
    assert np.all(w[4 + 2 * nd] == bz) and np.all(w[5 + 2 * nd] == 0.0) and np.all(w[6 + 2 * nd] == 0.0)
    assert np.all(w[:npert(nd)] == 0.0)

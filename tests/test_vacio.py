"""VAC configuration files (.ini ASCII, .out binary) through the C ABI.

The .out layout is checked by an independent numpy parser written from the
format description (src/readwrite.c:160-366, python/sacio.py read_sac_bin);
the .ini reader is checked on the reference's own input fixtures
(configs/*.ini) when the reference tree is mounted, and by round trips.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from smaug_amd import engine, vacio
from smaug_amd.params import nvar

REF_CONFIGS = "/root/reference/smaug/configs"

pytestmark = pytest.mark.skipif(not os.path.exists(engine.LIB_PATH), reason="libsmaughip.so not built")


def parse_out(path):
    """Independent reader of the binary snapshot."""
    with open(path, "rb") as fh:
        raw = fh.read()
    off = 0
    name = raw[off:off + 79]
    off += 79
    it = int(np.frombuffer(raw, np.int32, 1, off)[0])
    off += 4
    t = float(np.frombuffer(raw, np.float64, 1, off)[0])
    off += 8
    ndim, neqpar, nw = (int(v) for v in np.frombuffer(raw, np.int32, 3, off))
    off += 12
    n = [int(v) for v in np.frombuffer(raw, np.int32, ndim, off)]
    off += 4 * ndim
    eqpar = np.frombuffer(raw, np.float64, neqpar, off).copy()
    off += 8 * neqpar
    names = raw[off:off + 79]
    off += 79
    N = int(np.prod(n))
    data = np.frombuffer(raw, np.float64, (ndim + nw) * N, off).reshape(ndim + nw, N).copy()
    off += 8 * (ndim + nw) * N
    tail = raw[off:]
    return dict(name=name, it=it, t=t, ndim=ndim, neqpar=neqpar, nw=nw, n=n, eqpar=eqpar, names=names,
                pos=data[:ndim], w=data[ndim:], tail=tail)


@pytest.mark.parametrize("ndim,n", [(2, (12, 9)), (3, (7, 6, 5))])
def test_out_layout(tmp_path, ndim, n):
    rng = np.random.default_rng(1)
    N = int(np.prod(n))
    w = rng.standard_normal((nvar(ndim), N))
    pos = rng.standard_normal((ndim, N))
    path = str(tmp_path / "snap_7.out")
    g = (-274.0, 1.5, 2.5)[:ndim] + (0.0,) * (3 - ndim)
    vacio.write_out(path, ndim, n, w, pos, it=7, t=12.5, gamma=5 / 3, eta=0.25, g=g, name="fluxtube")
    d = parse_out(path)
    assert d["name"].rstrip(b"\0") == b"fluxtube"
    assert (d["it"], d["t"], d["ndim"], d["nw"]) == (7, 12.5, ndim, nvar(ndim))
    assert d["neqpar"] == (7 if ndim == 3 else 6)
    assert d["n"] == list(n)
    expect_eq = [5 / 3, 0.25, g[0], g[1]] + ([g[2]] if ndim == 3 else []) + [0.0, 0.0]
    assert np.array_equal(d["eqpar"], expect_eq)
    want = b"x y z rho mx my mz e bx by bz gamma eta g1 g2 g3" if ndim == 3 else b"x y rho mx my e bx by  gamma eta g1 g2"
    assert d["names"].rstrip(b"\0") == want
    assert np.array_equal(d["pos"], pos) and np.array_equal(d["w"], w)
    assert d["tail"] == b"\n"


@pytest.mark.parametrize("ndim,n", [(2, (12, 9)), (3, (7, 6, 5))])
def test_ini_round_trip(tmp_path, ndim, n):
    rng = np.random.default_rng(2)
    N = int(np.prod(n))
    w = rng.standard_normal((nvar(ndim), N)) * 10.0 ** rng.integers(-8, 8, (nvar(ndim), N))
    pos = rng.uniform(0, 4e6, (ndim, N))
    path = str(tmp_path / "cfg.ini")
    vacio.write_ini(path, ndim, n, w, pos, it=3, t=0.5, name="opozmf_mhd22")
    w2, pos2, it, t = vacio.read_ini(path, ndim, n)
    assert (it, t) == (3, 0.5)
    # rows are written with %G: 6 significant digits, as the reference does
    assert np.allclose(w2, w, rtol=5e-6, atol=0) and np.allclose(pos2, pos, rtol=5e-6, atol=0)
    with open(path) as fh:
        head = [next(fh) for _ in range(5)]
    assert head[0].strip() == "opozmf_mhd22"
    assert head[1].split()[2:] == [str(ndim), "7" if ndim == 3 else "6", str(nvar(ndim))]


def test_ini_rejects_other_grids(tmp_path):
    rng = np.random.default_rng(3)
    path = str(tmp_path / "cfg.ini")
    vacio.write_ini(path, 2, (8, 8), rng.standard_normal((10, 64)), rng.standard_normal((2, 64)))
    with pytest.raises(engine.SmaugError, match="grid dim"):
        vacio.read_ini(path, 2, (8, 9))
    with pytest.raises(engine.SmaugError, match="ndim"):
        vacio.read_ini(path, 3, (8, 8, 1))
    with pytest.raises(engine.SmaugError, match="cannot open"):
        vacio.read_ini(str(tmp_path / "absent.ini"), 2, (8, 8))


@pytest.mark.skipif(not os.path.isdir(REF_CONFIGS), reason="reference input fixtures not mounted")
def test_reads_reference_brio_wu_fixture():
    """configs/zero1_BW_asc.ini: 800 x 6 cells, Brio-Wu left state in row 1."""
    w, pos, it, t = vacio.read_ini(os.path.join(REF_CONFIGS, "zero1_BW_asc.ini"), 2, (800, 6))
    assert (it, t) == (0, 0.0)
    assert pos[0, 0] == 6.25e-4 and pos[1, 0] == pytest.approx(8.3333333333e-2, rel=1e-12)
    # rho, m1, m2, e, b1, b2, eb, rhob, b1b, b2b of the first cell
    assert np.array_equal(w[:, 0], [1.0, 0.0, 0.0, 1.78125, 0.75, 1.0, 0.0, 0.0, 0.0, 0.0])
    assert np.all(w[0] > 0)


@pytest.mark.skipif(not os.path.isdir(REF_CONFIGS), reason="reference input fixtures not mounted")
def test_reads_reference_orszag_tang_pieces():
    """configs/zero1_ot_asc_252_np0202_00{0..3}.ini: four 126 x 126 pieces of a
    252^2 Orszag-Tang start; uniform density and pressure-consistent energy."""
    for k in range(4):
        w, pos, it, t = vacio.read_ini(os.path.join(REF_CONFIGS, f"zero1_ot_asc_252_np0202_00{k}.ini"), 2, (126, 126))
        assert it == 0 and t == 0.0
        assert np.allclose(w[0], 2.2104852592e-01, rtol=1e-10)
        assert np.all(w[6:] == 0.0)


@pytest.mark.gpu
def test_engine_snapshot_matches_state(tmp_path):
    """smaug_write_out writes the context's current state and solver grid."""
    from smaug_amd.engine import Engine
    from cases import strat_case

    rp, w0, pos0 = strat_case(ni=(10, 8, 8), rkon=1)
    print("[vacio] engine snapshot, strat 10x8x8 RK4 x2", flush=True)
    with Engine(rp, w0, pos0) as e:
        e.step(2)
        path = e.write_out(str(tmp_path / "strat"), "fluxtube")
        state = e.state()
        pos, _ = e.grid()
        it, t, _, _ = e.clock()
    d = parse_out(path)
    assert path.endswith("strat_2.out") and d["it"] == it == 2 and d["t"] == t
    assert np.array_equal(d["w"], state) and np.array_equal(d["pos"], pos)

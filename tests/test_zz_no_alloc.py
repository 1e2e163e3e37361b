"""smaug_cfl_noalloc (the entry point that allocates nothing) against
smaug_cfl, the entry point it is a variant of: the per-direction wave-speed
maxima and the CFL step must be bitwise equal, on consecutive calls whose
grids shrink (the shared scratch keeps the last call's contents, so each
call must clear it), and the blow-up flag must be raised by a non-finite
state and cleared by the next clean one.
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from tests.cases import strat_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# (interior cells, ndim, moddton, NaN planted): the grids shrink call by call
SEQUENCE = [((24, 20, 16), 3, 1, False), ((20, 12), 2, 0, False), ((16, 12), 2, 1, True), ((12, 10), 2, 1, False)]


def reference(ni, nd, moddton, nan):
    """smaug_cfl of a fresh context on the same state (None for the planted
    NaN, which smaug_cfl reports as an error) and that context's state."""
    from smaug_amd.engine import Engine, SmaugError

    rp, w, pos = strat_case(ni=ni, ndim=nd, rkon=1)
    rp.moddton = moddton
    rp.courant = 0.3
    if moddton:
        rp.dt = 10.0  # above the CFL step, so the rule picks courant / courantmax
    if nan:
        w = w.copy()
        w[1 + nd, 5 * rp.n[0] + 6] = np.nan  # energy of one interior cell
    e = Engine(rp, w, pos)
    state = e.state()  # before smaug_cfl, which raises the blow-up flag on the NaN
    try:
        cm, dt = e.cfl()
        ref = (np.asarray(cm, dtype=np.float64), float(dt))
    except SmaugError:
        ref = None
    e.close()
    return rp, state, ref


@pytest.mark.gpu
def test_cfl_noalloc_matches_cfl_on_gpu():
    torch = pytest.importorskip("torch")
    from smaug_amd.engine import cfl_noalloc

    assert torch.cuda.is_available()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    for ni, nd, moddton, nan in SEQUENCE:
        print(f"[noalloc] {nd}D {ni} moddton={moddton} nan={nan}", flush=True)
        rp, state, ref = reference(ni, nd, moddton, nan)
        wd = torch.tensor(state.ravel(), dtype=torch.float64, device=dev)
        out = torch.full((nd + 2,), float("nan"), dtype=torch.float64, device=dev)
        cfl_noalloc(rp, wd.data_ptr(), out.data_ptr(), stream.cuda_stream)
        stream.synchronize()
        got = out.cpu().numpy()
        print(f"  got {got.tolist()} ref {None if ref is None else (ref[0].tolist(), ref[1])}", flush=True)
        if nan:
            assert ref is None and got[nd + 1] == 1.0
        else:
            assert got[nd + 1] == 0.0
            assert np.array_equal(got[:nd], ref[0]), (got, ref)
            assert got[nd] == ref[1], (got, ref)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_cfl_noalloc_matches_cfl_in_host_build():
    """The same comparison on the kernels compiled for the host (no GPU),
    with the states in guard-paged caller buffers."""
    sys.path.insert(0, ROOT)
    from tests.hostemu.build import build

    lib = build()
    env = dict(os.environ, SMAUG_HIP_LIB=lib, OMP_NUM_THREADS="1")
    p = subprocess.run([sys.executable, "-m", "tests.hostemu.check", "noalloc"], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, f"host check died ({p.returncode}):\n{p.stdout[-2000:]}\n{p.stderr[-4000:]}"
    for r in json.loads(p.stdout.strip().splitlines()[-1]):
        nd = r["ndim"]
        assert r["got"][:nd] == r["cmax"], r
        assert r["got"][nd] == r["dt"] and r["got"][nd + 1] == 0.0, r

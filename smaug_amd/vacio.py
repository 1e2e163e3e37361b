"""VAC configuration files through the C ABI (``smaug_ini_read``,
``smaug_ini_write``, ``smaug_out_write`` in libsmaughip.so; host code, no
device needed).

Formats are the reference's (src/readwrite.c): the ASCII ``.ini`` (five
header lines, then one row per cell, i fastest, ghosts included: positions
then the NVAR fields) and the binary ``.out`` snapshot (79-byte name, it, t,
ndim, neqpar, nw, n[], eqpar[], 79-byte variable names, position and field
blocks of float64, trailing newline).
"""
from __future__ import annotations

import ctypes
from typing import Sequence, Tuple

import numpy as np

from .engine import _chk, _dp, load
from .params import nvar


def _ip(n: Sequence[int]):
    arr = (ctypes.c_int * 3)(*(list(n) + [1] * (3 - len(n))))
    return arr


def read_ini(path: str, ndim: int, n: Sequence[int]) -> Tuple[np.ndarray, np.ndarray, int, float]:
    """Returns (w[NVAR, N], pos[ndim, N], it, t) of an ASCII .ini file."""
    N = int(np.prod(n[:ndim]))
    w = np.empty((nvar(ndim), N))
    pos = np.empty((ndim, N))
    it, t = ctypes.c_int(), ctypes.c_double()
    _chk(load().smaug_ini_read(path.encode(), ndim, _ip(n), _dp(w), _dp(pos), ctypes.byref(it), ctypes.byref(t)),
         "smaug_ini_read")
    return w, pos, it.value, t.value


def _write(fn, what, path, name, ndim, n, w, pos, it, t, gamma, eta, g):
    w = np.ascontiguousarray(w, dtype=np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    gg = (ctypes.c_double * 3)(*(list(g) + [0.0] * (3 - len(g)))[:3])
    _chk(fn(path.encode(), name.encode(), ndim, _ip(n), _dp(w), _dp(pos), int(it), float(t), float(gamma),
            float(eta), gg), what)


def write_ini(path, ndim, n, w, pos, it=0, t=0.0, gamma=5.0 / 3.0, eta=0.0, g=(0.0, 0.0, 0.0), name=""):
    _write(load().smaug_ini_write, "smaug_ini_write", path, name, ndim, n, w, pos, it, t, gamma, eta, g)


def write_out(path, ndim, n, w, pos, it=0, t=0.0, gamma=5.0 / 3.0, eta=0.0, g=(0.0, 0.0, 0.0), name=""):
    _write(load().smaug_out_write, "smaug_out_write", path, name, ndim, n, w, pos, it, t, gamma, eta, g)

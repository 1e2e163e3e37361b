"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; never by the product.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.environ.get("SMAUG_ORACLE_LIB", os.path.join(_HERE, "liboracle.so"))
_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        L.orc_create.argtypes = [ctypes.c_void_p, dp, dp, ctypes.POINTER(vp)]
        L.orc_step.argtypes = [vp, ctypes.c_int]
        L.orc_get_state.argtypes = [vp, dp]
        L.orc_set_state.argtypes = [vp, dp]
        L.orc_get_wd.argtypes = [vp, dp]
        L.orc_cfl.argtypes = [vp, dp, dp]
        L.orc_get_clock.argtypes = [vp, ip, dp, dp, dp]
        L.orc_destroy.argtypes = [vp]
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_probe_grad.argtypes = [vp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, dp]
        L.orc_probe_visc.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, dp, dp]
        L.orc_set_driver_map.argtypes = [vp, dp]
        L.orc_set_driver_param.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.orc_restart.argtypes = [vp, dp, ctypes.c_int, ctypes.c_double]
        L.orc_nu_mode.argtypes = [vp, ctypes.c_int]
        L.orc_nu_mode.restype = ctypes.c_long
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """CPU restatement of the reference step on a copy of (w0, pos)."""

    def __init__(self, rp, w0: np.ndarray, pos: np.ndarray):
        self.rp = rp
        self._cp = rp.to_c()
        self.nd = rp.ndim
        self.N = rp.ncells
        self.nvar = 4 + 3 * rp.ndim
        w0 = np.ascontiguousarray(w0, dtype=np.float64).reshape(self.nvar, self.N)
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.nd, self.N)
        h = ctypes.c_void_p()
        rc = lib().orc_create(ctypes.byref(self._cp), _dp(w0), _dp(pos), ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"orc_create failed ({rc})")
        self._h = h

    def step(self, n: int = 1):
        lib().orc_step(self._h, int(n))

    def state(self) -> np.ndarray:
        w = np.empty((self.nvar, self.N))
        lib().orc_get_state(self._h, _dp(w))
        return w

    def set_state(self, w: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float64)
        lib().orc_set_state(self._h, _dp(w))

    def wd(self) -> np.ndarray:
        nderv = 4 * self.nd + 11
        out = np.empty((nderv, self.N))
        lib().orc_get_wd(self._h, _dp(out))
        return out

    def grid(self):
        wd = self.wd()
        nd = self.nd
        return wd[2 * nd + 11: 3 * nd + 11].copy(), wd[3 * nd + 11: 4 * nd + 11].copy()

    def cfl(self):
        cm = np.zeros(3)
        dt = ctypes.c_double()
        lib().orc_cfl(self._h, _dp(cm), ctypes.byref(dt))
        return cm[: self.nd], dt.value

    def clock(self):
        it, t, qt, dt = ctypes.c_int(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        lib().orc_get_clock(self._h, ctypes.byref(it), ctypes.byref(t), ctypes.byref(qt), ctypes.byref(dt))
        return it.value, t.value, qt.value, dt.value

    GRAD_KINDS = {"grad3dn": 0, "grad1l3n": 1, "grad1r3n": 2, "grad13n": 3}

    def grad(self, kind: str, a: np.ndarray, bcidx: int, direction: int) -> np.ndarray:
        """One reference differential operator applied to ``a`` at every cell."""
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.N)
        out = np.empty(self.N)
        rc = lib().orc_probe_grad(self._h, self.GRAD_KINDS[kind], _dp(a), int(bcidx), int(direction), _dp(out))
        if rc != 0:
            raise ValueError("orc_probe_grad rejected its arguments")
        return out

    def visc(self, field: int, dim: int):
        """(nu_r, nu_l, cmax) of one field of the current state along dim."""
        nur, nul = np.empty(self.N), np.empty(self.N)
        cm = ctypes.c_double()
        rc = lib().orc_probe_visc(self._h, int(field), int(dim), _dp(nur), _dp(nul), ctypes.byref(cm))
        if rc != 0:
            raise ValueError("orc_probe_visc rejected its arguments")
        return nur, nul, cm.value

    def set_driver_map(self, amap):
        """Driver amplitude per (j, k) column (n1*n2 values), None clears."""
        if amap is None:
            lib().orc_set_driver_map(self._h, None)
            return
        a = np.ascontiguousarray(amap, dtype=np.float64).reshape(-1)
        n = self.rp.n
        if a.size != n[1] * (n[2] if self.nd == 3 else 1):
            raise ValueError("driver map needs one value per (j, k) column")
        lib().orc_set_driver_map(self._h, _dp(a))

    def set_driver_param(self, q: int, v: float):
        """Driver parameter drv[q] for the steps that follow."""
        if lib().orc_set_driver_param(self._h, int(q), float(v)) != 0:
            raise ValueError("driver parameter index out of range")

    def restart(self, w: np.ndarray, it: int = 0, qt: float = 0.0):
        """Start again from state w (nvar x N) at step it and clock qt."""
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(self.nvar, self.N)
        lib().orc_restart(self._h, _dp(w), int(it), float(qt))

    NU_OFF, NU_RECORD, NU_REPLAY = 0, 1, 2

    def nu_mode(self, mode: int) -> int:
        """Hyperdiffusion coefficients: 0 computed, 1 computed and recorded,
        2 replayed from the recording (frozen coefficients). Returns the
        number of recorded coefficient pairs."""
        return int(lib().orc_nu_mode(self._h, int(mode)))

    def close(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def num_threads() -> int:
    return lib().orc_num_threads()

"""ctypes binding of the C-ABI (include/smaug_hip.h) and the host-side
driver loop.

The engine runs only through the HIP library ``libsmaughip.so``; there is no
CPU fallback: a missing library or a failed device call raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Tuple

import numpy as np

from .params import RunParams, nvar

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SMAUG_HIP_LIB", os.path.join(_HERE, "libsmaughip.so"))
_lib = None

EXPORTS = [
    "smaug_create", "smaug_step", "smaug_cfl", "smaug_fill_ghosts", "smaug_get_state",
    "smaug_set_state", "smaug_get_grid", "smaug_get_clock", "smaug_state_device_ptr",
    "smaug_sync", "smaug_destroy", "smaug_last_error", "smaug_version", "smaug_time_kernels",
    "smaug_ini_read", "smaug_ini_write", "smaug_out_write", "smaug_write_out",
    "smaug_set_driver_map", "smaug_set_stream", "smaug_set_clock", "smaug_save_pert", "smaug_load_pert",
    "smaug_tape_doubles", "smaug_step_taped", "smaug_adjoint_step", "smaug_cfl_noalloc",
    "smaug_tl_begin", "smaug_tl_driver_amplitude", "smaug_tl_step", "smaug_tl_get", "smaug_tl_end",
]


class SmaugError(RuntimeError):
    pass


def load() -> ctypes.CDLL:
    """Loads libsmaughip.so (built by smaug_amd/build.py); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SmaugError(f"{LIB_PATH} is missing: build it with `python -m smaug_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    L.smaug_create.argtypes = [vp, dp, dp, ctypes.POINTER(vp)]
    L.smaug_step.argtypes = [vp, ctypes.c_int]
    L.smaug_cfl.argtypes = [vp, dp, dp]
    L.smaug_fill_ghosts.argtypes = [vp]
    L.smaug_get_state.argtypes = [vp, dp]
    L.smaug_set_state.argtypes = [vp, dp]
    L.smaug_get_grid.argtypes = [vp, dp, dp]
    L.smaug_get_clock.argtypes = [vp, ip, dp, dp, dp]
    L.smaug_state_device_ptr.argtypes = [vp, ctypes.POINTER(vp)]
    L.smaug_sync.argtypes = [vp]
    L.smaug_destroy.argtypes = [vp]
    L.smaug_last_error.restype = ctypes.c_char_p
    L.smaug_version.restype = ctypes.c_char_p
    L.smaug_time_kernels.argtypes = [vp, ctypes.c_int, dp, dp, dp]
    cp = ctypes.c_char_p
    L.smaug_ini_read.argtypes = [cp, ctypes.c_int, ip, dp, dp, ip, dp]
    L.smaug_ini_write.argtypes = [cp, cp, ctypes.c_int, ip, dp, dp, ctypes.c_int, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, dp]
    L.smaug_out_write.argtypes = L.smaug_ini_write.argtypes
    L.smaug_write_out.argtypes = [vp, cp, cp]
    L.smaug_set_driver_map.argtypes = [vp, vp, ctypes.c_int]
    L.smaug_set_stream.argtypes = [vp, vp]
    L.smaug_set_clock.argtypes = [vp, ctypes.c_int, ctypes.c_double]
    L.smaug_save_pert.argtypes = [vp, vp]
    L.smaug_load_pert.argtypes = [vp, vp]
    L.smaug_tape_doubles.argtypes = [vp]
    L.smaug_step_taped.argtypes = [vp, vp]
    L.smaug_adjoint_step.argtypes = [vp, vp, vp, vp, vp]
    L.smaug_tl_begin.argtypes = [vp, dp, dp, dp]
    L.smaug_tl_driver_amplitude.argtypes = [vp, ctypes.c_double]
    L.smaug_tl_step.argtypes = [vp, ctypes.c_int]
    L.smaug_tl_get.argtypes = [vp, dp, dp]
    L.smaug_tl_end.argtypes = [vp]
    L.smaug_cfl_noalloc.argtypes = [vp, vp, vp, vp]
    for name in EXPORTS:
        getattr(L, name).restype = ctypes.c_int
    L.smaug_last_error.restype = ctypes.c_char_p
    L.smaug_version.restype = ctypes.c_char_p
    L.smaug_tape_doubles.restype = ctypes.c_longlong
    _lib = L
    return L


def _chk(rc: int, what: str):
    if rc != 0:
        msg = load().smaug_last_error().decode(errors="replace")
        raise SmaugError(f"{what} failed ({rc}): {msg}")


def cfl_noalloc(rp: RunParams, w_ptr: int, out_ptr: int, stream_ptr: int = 0):
    """smaug_cfl_noalloc: the wave-speed maxima per direction, the CFL step
    and the blow-up flag of a device state (NVAR*N doubles at w_ptr, the
    layout of Engine's w0) into ndim + 2 doubles at out_ptr, queued on the
    HIP stream stream_ptr (0: the null stream) without allocating anything.
    Read out_ptr after synchronising that stream."""
    cp = rp.to_c()
    _chk(load().smaug_cfl_noalloc(ctypes.byref(cp), ctypes.c_void_p(w_ptr), ctypes.c_void_p(out_ptr),
                                  ctypes.c_void_p(stream_ptr or None)), "smaug_cfl_noalloc")


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Engine:
    """One simulation on the GPU: state, grid and the time loop."""

    def __init__(self, rp: RunParams, w0: np.ndarray, pos: np.ndarray):
        L = load()
        self.rp = rp
        self._cp = rp.to_c()
        self.nd = rp.ndim
        self.N = rp.ncells
        self.nvar = nvar(rp.ndim)
        w0 = np.ascontiguousarray(w0, dtype=np.float64).reshape(self.nvar, self.N)
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.nd, self.N)
        h = ctypes.c_void_p()
        _chk(L.smaug_create(ctypes.byref(self._cp), _dp(w0), _dp(pos), ctypes.byref(h)), "smaug_create")
        self._h = h

    def step(self, n: int = 1):
        _chk(load().smaug_step(self._h, int(n)), "smaug_step")

    def sync(self):
        _chk(load().smaug_sync(self._h), "smaug_sync")

    def state(self) -> np.ndarray:
        w = np.empty((self.nvar, self.N))
        _chk(load().smaug_get_state(self._h, _dp(w)), "smaug_get_state")
        return w

    def set_state(self, w: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float64)
        _chk(load().smaug_set_state(self._h, _dp(w)), "smaug_set_state")

    def fill_ghosts(self):
        _chk(load().smaug_fill_ghosts(self._h), "smaug_fill_ghosts")

    def grid(self) -> Tuple[np.ndarray, np.ndarray]:
        pos = np.empty((self.nd, self.N))
        dx = np.empty((self.nd, self.N))
        _chk(load().smaug_get_grid(self._h, _dp(pos), _dp(dx)), "smaug_get_grid")
        return pos, dx

    def cfl(self) -> Tuple[np.ndarray, float]:
        cm = np.zeros(3)
        dt = ctypes.c_double()
        _chk(load().smaug_cfl(self._h, _dp(cm), ctypes.byref(dt)), "smaug_cfl")
        return cm[: self.nd], dt.value

    def clock(self):
        it, t, qt, dt = ctypes.c_int(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _chk(load().smaug_get_clock(self._h, ctypes.byref(it), ctypes.byref(t), ctypes.byref(qt),
                                    ctypes.byref(dt)), "smaug_get_clock")
        return it.value, t.value, qt.value, dt.value

    def device_ptr(self) -> int:
        p = ctypes.c_void_p()
        _chk(load().smaug_state_device_ptr(self._h, ctypes.byref(p)), "smaug_state_device_ptr")
        return p.value

    def write_out(self, prefix: str, name: str = "") -> str:
        """Binary VAC snapshot of the current state: <prefix>_<it>.out."""
        _chk(load().smaug_write_out(self._h, prefix.encode(), name.encode()), "smaug_write_out")
        return f"{prefix}_{self.clock()[0]}.out"

    # ---- device-pointer entry points (buffers owned by the caller, e.g. torch)
    @property
    def npert(self) -> int:
        return 2 + 2 * self.nd

    def set_driver_map(self, amap=None, device_ptr: int = 0):
        """Driver amplitude per (j, k) column: a host array, or a device
        pointer (device_ptr) to n1*n2 doubles; None and 0 clear it."""
        L = load()
        if device_ptr:
            _chk(L.smaug_set_driver_map(self._h, ctypes.c_void_p(device_ptr), 1), "smaug_set_driver_map")
        elif amap is None:
            _chk(L.smaug_set_driver_map(self._h, None, 0), "smaug_set_driver_map")
        else:
            a = np.ascontiguousarray(amap, dtype=np.float64).reshape(-1)
            if a.size != self.N // self.rp.n[0]:
                raise ValueError("driver map needs one value per (j, k) column")
            _chk(L.smaug_set_driver_map(self._h, a.ctypes.data_as(ctypes.c_void_p), 0), "smaug_set_driver_map")

    def set_stream(self, stream_ptr: int = 0):
        _chk(load().smaug_set_stream(self._h, ctypes.c_void_p(stream_ptr or None)), "smaug_set_stream")

    def set_clock(self, it: int, qt: float):
        _chk(load().smaug_set_clock(self._h, int(it), float(qt)), "smaug_set_clock")

    def save_pert(self, dst_ptr: int):
        _chk(load().smaug_save_pert(self._h, ctypes.c_void_p(dst_ptr)), "smaug_save_pert")

    def load_pert(self, src_ptr: int):
        _chk(load().smaug_load_pert(self._h, ctypes.c_void_p(src_ptr)), "smaug_load_pert")

    def tape_doubles(self) -> int:
        n = int(load().smaug_tape_doubles(self._h))
        if n < 0:
            _chk(n, "smaug_tape_doubles")
        return n

    def step_taped(self, tape_ptr: int):
        _chk(load().smaug_step_taped(self._h, ctypes.c_void_p(tape_ptr)), "smaug_step_taped")

    def adjoint_step(self, tape_ptr: int, wnext_ptr: int, lam_ptr: int, gmap_ptr: int = 0):
        _chk(load().smaug_adjoint_step(self._h, ctypes.c_void_p(tape_ptr), ctypes.c_void_p(wnext_ptr),
                                       ctypes.c_void_p(lam_ptr), ctypes.c_void_p(gmap_ptr or None)),
             "smaug_adjoint_step")

    # ---- tangent-linear run (forward-mode derivative; host arrays)
    def tl_begin(self, dw0: np.ndarray, dbg: Optional[np.ndarray] = None, damap: Optional[np.ndarray] = None):
        """Starts a tangent-linear run from the current state in the direction
        dw0 (np x N evolved fields), dbg (nb x N background fields e_b, rho_b,
        b_b; None = 0) and, with a driver map set, damap (one value per
        (j, k) column)."""
        a = np.ascontiguousarray(dw0, dtype=np.float64).reshape(-1)
        if a.size != self.npert * self.N:
            raise ValueError("dw0 needs np*N values")
        b = None
        if dbg is not None:
            b = np.ascontiguousarray(dbg, dtype=np.float64).reshape(-1)
            if b.size != (self.nvar - self.npert) * self.N:
                raise ValueError("dbg needs nb*N values")
        m = None
        if damap is not None:
            m = np.ascontiguousarray(damap, dtype=np.float64).reshape(-1)
            if m.size != self.N // self.rp.n[0]:
                raise ValueError("damap needs one value per (j, k) column")
        _chk(load().smaug_tl_begin(self._h, _dp(a), None if b is None else _dp(b), None if m is None else _dp(m)),
             "smaug_tl_begin")

    def tl_driver_amplitude(self, d: float):
        """Tangent of the driver amplitude drv[0] in the current run's direction."""
        _chk(load().smaug_tl_driver_amplitude(self._h, float(d)), "smaug_tl_driver_amplitude")

    def tl_step(self, n: int = 1):
        _chk(load().smaug_tl_step(self._h, int(n)), "smaug_tl_step")

    def tl_get(self) -> Tuple[np.ndarray, np.ndarray]:
        """(state, tangent) of the run, each np x N."""
        w = np.empty((self.npert, self.N))
        dw = np.empty((self.npert, self.N))
        _chk(load().smaug_tl_get(self._h, _dp(w), _dp(dw)), "smaug_tl_get")
        return w, dw

    def tl_end(self):
        _chk(load().smaug_tl_end(self._h), "smaug_tl_end")

    def time_kernels(self, nsteps: int) -> Tuple[float, float, float]:
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _chk(load().smaug_time_kernels(self._h, int(nsteps), ctypes.byref(a), ctypes.byref(b),
                                       ctypes.byref(c)), "smaug_time_kernels")
        return a.value, b.value, c.value

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_torch_stream", None) is not None:
                load().smaug_set_stream(self._h, None)  # back to the own stream first
            load().smaug_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

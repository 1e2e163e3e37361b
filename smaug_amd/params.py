"""Run parameters of the SMAUG time step, mirrored from the reference.

The reference fixes its parameters at compile time with ``#define``s in
``include/iosmaugparams.h`` (one file per model under
``models/iosmaugparams.h.*``) and copies them into ``struct params`` in
``src/iosmaug.c:91-232``.  Here the same names are read from a runtime
``key = value`` file (or a dict) and packed into the C-ABI struct
``smaug_params`` declared in ``include/smaug_hip.h``.

Recognised keys (reference names): NI NJ NK (interior cells), XMIN XMAX
YMIN YMAX ZMIN ZMAX, DT, PGAMMA, PMU, PETA, PGRAV0..2, PCOURANT, PRKON,
PMODDTON, PHYPERDIFMOM, PDIVBON, PCHYP, PCHYPRHO, PCHYPENERGY, PCHYPB0..2,
PCHYPMOM0..2, BCU0..2, BCL0..2, plus DRIVER and DRV0..DRV11 for the user
source plug-in, PRECISION (fp64|mixed) and NDIM.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional

C = ctypes

NGHOST = 2  # include/iosmaugparams.h ngi/ngj/ngk

# boundary codes, include/iotypes.h:144
BC_PERIODIC, BC_MPI, BC_MPIPERIOD, BC_CONT, BC_CONTCD4, BC_FIXED, BC_SYMM, BC_ASYMM = range(8)

TERMS_ALL = 0xFF
T_FLUX, T_ETERM, T_GRAV, T_HRHO, T_HE, T_HMOM, T_HB, T_SOURCE = (1 << i for i in range(8))

DRIVER_NONE = 0
DRIVER_VERTICAL = 1
DRIVER_TORSIONAL = 2  # models/usersource.cu.fluxtube1
DRIVER_SPICULE1 = 3   # models/usersource.cu.spicule1
NDRV = 12

PREC_FP64 = 0
PREC_MIXED = 1


class SmaugParams(C.Structure):
    """ctypes image of ``smaug_params`` (include/smaug_hip.h)."""

    _fields_ = [
        ("ndim", C.c_int),
        ("n", C.c_int * 3),
        ("gamma", C.c_double),
        ("mu", C.c_double),
        ("eta", C.c_double),
        ("g", C.c_double * 3),
        ("dt", C.c_double),
        ("courant", C.c_double),
        ("rkon", C.c_int),
        ("moddton", C.c_int),
        ("hyperdifmom", C.c_int),
        ("divbon", C.c_int),
        ("chyp", C.c_double * 13),
        ("xmin", C.c_double * 3),
        ("xmax", C.c_double * 3),
        ("dx", C.c_double * 3),
        ("boundtype", ((C.c_int * 2) * 3) * 13),
        ("driver", C.c_int),
        ("drv", C.c_double * NDRV),
        ("terms", C.c_uint),
        ("precision", C.c_int),
        ("qt0", C.c_double),
        ("it0", C.c_int),
        ("divbfix", C.c_int),
    ]


def field_names(ndim: int) -> List[str]:
    """Variable order of the state vector (include/iotypes.h:211,215)."""
    if ndim == 2:
        return ["rho", "mom1", "mom2", "energy", "b1", "b2", "energyb", "rhob", "b1b", "b2b"]
    return ["rho", "mom1", "mom2", "mom3", "energy", "b1", "b2", "b3",
            "energyb", "rhob", "b1b", "b2b", "b3b"]


def nvar(ndim: int) -> int:
    return 4 + 3 * ndim


def npert(ndim: int) -> int:
    """Number of evolved (perturbation) fields: rho, mom, energy, b."""
    return 2 + 2 * ndim


@dataclass
class RunParams:
    """Parameter set of one run, with the reference's names and meanings."""

    ndim: int
    ni: List[int]                      # interior cells per dim (NI, NJ, NK)
    xmin: List[float]
    xmax: List[float]
    dt: float = 1.0                    # DT
    gamma: float = 5.0 / 3.0           # PGAMMA
    mu: float = 1.0                    # PMU
    eta: float = 0.0                   # PETA
    g: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    courant: float = 0.15              # PCOURANT
    rkon: int = 0                      # PRKON
    moddton: int = 1                   # PMODDTON
    hyperdifmom: int = 1               # PHYPERDIFMOM
    divbon: int = 0                    # PDIVBON
    divbfix: int = 0                   # PDIVBFIX
    chyp: Dict[str, float] = field(default_factory=dict)
    bcu: List[int] = field(default_factory=lambda: [0, 0, 0])   # BCU0..2
    bcl: List[int] = field(default_factory=lambda: [0, 0, 0])   # BCL0..2
    # per-field boundary types per dim, overriding bcu/bcl for the named
    # fields (the reference's boundtype[field][dir] table, include/iotypes.h:144)
    bc_field: Dict[str, List[int]] = field(default_factory=dict)
    driver: int = DRIVER_NONE
    drv: List[float] = field(default_factory=lambda: [0.0] * NDRV)
    terms: int = TERMS_ALL
    precision: int = PREC_FP64
    qt0: float = 0.0
    it0: int = 0

    @property
    def n(self) -> List[int]:
        """Cells per dim including ghosts (src/iosmaug.c:47-57)."""
        return [self.ni[d] + 2 * NGHOST for d in range(self.ndim)]

    @property
    def ncells(self) -> int:
        return int(math.prod(self.n))

    @property
    def dx(self) -> List[float]:
        """p->dx = (xmax-xmin)/NI (src/iosmaug.c:49-56)."""
        return [(self.xmax[d] - self.xmin[d]) / self.ni[d] for d in range(self.ndim)]

    def to_c(self) -> SmaugParams:
        p = SmaugParams()
        p.ndim = self.ndim
        n = self.n + [1] * (3 - self.ndim)
        for d in range(3):
            p.n[d] = n[d]
        p.gamma, p.mu, p.eta = self.gamma, self.mu, self.eta
        for d in range(3):
            p.g[d] = self.g[d] if d < len(self.g) else 0.0
        p.dt = self.dt
        p.courant = self.courant
        p.rkon, p.moddton, p.hyperdifmom, p.divbon = self.rkon, self.moddton, self.hyperdifmom, self.divbon
        names = field_names(self.ndim)
        default = self.chyp.get("default", 0.02)  # PCHYP
        for f in range(13):
            p.chyp[f] = 0.0
        for f, nm in enumerate(names[: npert(self.ndim)]):
            p.chyp[f] = self.chyp.get(nm, default)
        for d in range(self.ndim):
            p.xmin[d], p.xmax[d] = self.xmin[d], self.xmax[d]
            p.dx[d] = self.dx[d]
        all_names = field_names(self.ndim)
        for f in range(13):
            over = self.bc_field.get(all_names[f]) if f < len(all_names) else None
            for d in range(3):
                p.boundtype[f][d][0] = (over[d] if over else self.bcu[d]) if d < self.ndim else 0
                p.boundtype[f][d][1] = (over[d] if over else self.bcl[d]) if d < self.ndim else 0
        p.driver = self.driver
        for i in range(NDRV):
            p.drv[i] = self.drv[i] if i < len(self.drv) else 0.0
        p.terms = self.terms
        p.precision = self.precision
        p.qt0 = self.qt0
        p.it0 = self.it0
        p.divbfix = self.divbfix
        return p


_CHYP_KEYS = {
    "PCHYPRHO": "rho", "PCHYPENERGY": "energy",
    "PCHYPB0": "b1", "PCHYPB1": "b2", "PCHYPB2": "b3",
    "PCHYPMOM0": "mom1", "PCHYPMOM1": "mom2", "PCHYPMOM2": "mom3",
}


def parse_param_text(text: str) -> Dict[str, str]:
    """Parses ``KEY = value`` lines (``#`` starts a comment)."""
    out: Dict[str, str] = {}
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        if "=" not in line:
            raise ValueError(f"parameter line without '=': {raw!r}")
        k, v = line.split("=", 1)
        out[k.strip().upper()] = v.strip()
    return out


def from_keys(kv: Dict[str, str]) -> RunParams:
    """Builds RunParams from reference-named keys (see module docstring)."""
    ndim = int(kv.get("NDIM", "3" if "NK" in kv else "2"))
    ni = [int(kv["NI"]), int(kv["NJ"])] + ([int(kv["NK"])] if ndim == 3 else [])
    lo = [float(kv.get(k, 0.0)) for k in ("XMIN", "YMIN", "ZMIN")][:ndim]
    hi = [float(kv.get(k, 1.0)) for k in ("XMAX", "YMAX", "ZMAX")][:ndim]
    rp = RunParams(ndim=ndim, ni=ni, xmin=lo, xmax=hi)
    if "DT" in kv:
        rp.dt = float(kv["DT"])
    for key, attr in (("PGAMMA", "gamma"), ("PMU", "mu"), ("PETA", "eta"), ("PCOURANT", "courant")):
        if key in kv:
            setattr(rp, attr, float(kv[key]))
    for key, attr in (("PRKON", "rkon"), ("PMODDTON", "moddton"), ("PHYPERDIFMOM", "hyperdifmom"),
                      ("PDIVBON", "divbon"), ("PDIVBFIX", "divbfix"), ("DRIVER", "driver")):
        if key in kv:
            setattr(rp, attr, int(float(kv[key])))
    rp.g = [float(kv.get(f"PGRAV{d}", 0.0)) for d in range(3)]
    rp.bcu = [int(kv.get(f"BCU{d}", 0)) for d in range(3)]
    rp.bcl = [int(kv.get(f"BCL{d}", 0)) for d in range(3)]
    chyp = {"default": float(kv.get("PCHYP", 0.02))}
    for key, nm in _CHYP_KEYS.items():
        if key in kv:
            chyp[nm] = float(kv[key])
    rp.chyp = chyp
    rp.drv = [float(kv.get(f"DRV{i}", 0.0)) for i in range(NDRV)]
    prec = kv.get("PRECISION", "fp64").lower()
    rp.precision = PREC_MIXED if prec == "mixed" else PREC_FP64
    return rp


def load_param_file(path: str) -> RunParams:
    with open(path, "r", encoding="utf-8") as fh:
        return from_keys(parse_param_text(fh.read()))


def validate(rp: RunParams) -> Optional[str]:
    """Returns an error message for parameter sets the engine rejects (the
    same rules as smaug_create's parameter check, smaug_host.hip)."""
    if rp.ndim not in (2, 3):
        return "ndim must be 2 or 3"
    if any(n < 4 for n in rp.ni):
        return "need at least 4 interior cells per dim"
    cells = 1
    for d in range(rp.ndim):
        cells *= rp.ni[d] + 4
    if cells >= 1 << 29:
        return f"{cells} cells: grids of 2^29 cells or more are not supported"
    if rp.divbon not in (0, 1):
        return "divbon must be 0 or 1"
    tables = [list(rp.bcu[: rp.ndim]), list(rp.bcl[: rp.ndim])]
    tables += [list(t[: rp.ndim]) for t in getattr(rp, "bc_field", {}).values()]
    for t in tables:
        for d, bt in enumerate(t):
            if bt == BC_PERIODIC and rp.ndim == 3:
                return "periodic boundaries are supported in 2D only"
            if bt in (1, 2):
                return f"boundary type {bt} in dir {d}: the MPI boundary types need a multi-domain run"
            if not 0 <= bt <= 7:
                return f"unknown boundary type {bt} in dir {d}"
    dv = list(rp.drv) + [0.0] * NDRV
    if rp.driver not in (DRIVER_NONE, DRIVER_VERTICAL, DRIVER_TORSIONAL, DRIVER_SPICULE1):
        return "unknown driver"
    if rp.driver == DRIVER_VERTICAL and 0.0 in (dv[1], dv[5], dv[6]):
        return "the vertical driver needs a period and widths (DRV1, DRV5, DRV6)"
    if rp.driver == DRIVER_TORSIONAL and (rp.ndim != 3 or 0.0 in (dv[5], dv[6], dv[7], dv[8], dv[9])):
        return "the torsional driver is 3D and needs widths and scales (DRV5..DRV9)"
    if rp.driver == DRIVER_SPICULE1 and 0.0 in (dv[5], dv[6]):
        return "the spicule driver needs widths (DRV5, DRV6)"
    return None

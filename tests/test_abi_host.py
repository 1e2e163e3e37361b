"""C-ABI and host-side checks that need no GPU.

* libsmaughip.so loads and exports every entry point include/smaug_hip.h
  declares (no compute call is made: there is no device here);
* the ctypes image of smaug_params matches the C struct layout;
* the reference-named parameter reader and its validation.
"""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

from smaug_amd import engine, params
from smaug_amd.params import BC_CONTCD4, BC_PERIODIC, PREC_MIXED, from_keys, parse_param_text, validate

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "smaug_hip.h")


def header_functions():
    text = open(HEADER, encoding="utf-8").read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(smaug_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert header_functions() == sorted(engine.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("libsmaughip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(engine.LIB_PATH)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing
    lib.smaug_version.restype = ctypes.c_char_p
    assert lib.smaug_version().decode().startswith("smaug-hip")


def test_library_exports_nothing_else_with_c_linkage():
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("libsmaughip.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", engine.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = sorted({ln.split()[-1] for ln in out.splitlines() if re.search(r" T smaug_", ln)})
    assert exported == header_functions()


def test_params_struct_layout_matches_c():
    """sizeof/offsets of smaug_params as compiled by the C compiler."""
    src = r'''
#include <stddef.h>
#include <stdio.h>
#include "smaug_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(smaug_params), offsetof(smaug_params, gamma),
         offsetof(smaug_params, chyp), offsetof(smaug_params, boundtype), offsetof(smaug_params, drv),
         offsetof(smaug_params, terms), offsetof(smaug_params, it0), offsetof(smaug_params, divbfix));
  return 0;
}
'''
    tmp = os.path.join(ROOT, "tests", "_layout_probe")
    try:
        with open(tmp + ".c", "w") as fh:
            fh.write(src)
        subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), tmp + ".c", "-o", tmp],
                       check=True)
        vals = [int(v) for v in subprocess.run([tmp], capture_output=True, text=True, check=True).stdout.split()]
    finally:
        for f in (tmp + ".c", tmp):
            if os.path.exists(f):
                os.remove(f)
    S = params.SmaugParams
    assert vals == [ctypes.sizeof(S), S.gamma.offset, S.chyp.offset, S.boundtype.offset, S.drv.offset,
                    S.terms.offset, S.it0.offset, S.divbfix.offset]


def test_param_text_uses_reference_names():
    text = """
    # fluxtube-like run
    NI = 32
    NJ = 16
    NK = 8
    XMIN = 1.0e5 ; XMAX = 6.0e6
    """.replace(" ; ", "\n")
    text += "YMAX = 4.0e6\nZMAX = 4.0e6\nDT = 0.5\nPGAMMA = 1.6666666666666667\nPGRAV0 = -274.0\n"
    text += "PCOURANT = 0.2\nPRKON = 1\nPMODDTON = 0\nPCHYP = 0.02\nPCHYPMOM0 = 0.2\n"
    text += "BCU0 = 4\nBCL0 = 4\nBCU1 = 4\nBCL1 = 4\nBCU2 = 4\nBCL2 = 4\nPRECISION = mixed\n"
    rp = from_keys(parse_param_text(text))
    assert rp.ndim == 3 and rp.ni == [32, 16, 8] and rp.n == [36, 20, 12]
    assert rp.dx[0] == pytest.approx((6.0e6 - 1.0e5) / 32)
    assert rp.g[0] == -274.0 and rp.rkon == 1 and rp.moddton == 0 and rp.dt == 0.5
    assert rp.precision == PREC_MIXED
    c = rp.to_c()
    assert c.chyp[1] == 0.2 and c.chyp[0] == 0.02 and c.chyp[2] == 0.02
    assert c.boundtype[4][2][1] == BC_CONTCD4
    assert validate(rp) is None


def test_param_text_rejects_malformed_lines():
    with pytest.raises(ValueError):
        parse_param_text("NI 32\n")


@pytest.mark.parametrize("mutate,needle", [
    (lambda rp: setattr(rp, "bcu", [BC_PERIODIC] * 3), "periodic"),
    (lambda rp: setattr(rp, "bcu", [1, 4, 4]), "boundary type 1"),
    (lambda rp: setattr(rp, "bcu", [9, 4, 4]), "unknown boundary type 9"),
    (lambda rp: setattr(rp, "ni", [3, 8, 8]), "4 interior"),
    (lambda rp: setattr(rp, "ni", [1020, 1020, 512]), "2^29"),
])
def test_validate_names_unsupported_settings(mutate, needle):
    from smaug_amd import configs
    rp = configs.stratified_params(ni=(8, 8, 8))
    mutate(rp)
    msg = validate(rp)
    assert msg is not None and needle in msg


@pytest.mark.parametrize("bt", [BC_CONTCD4, 3, 5, 6, 7])
def test_validate_accepts_engine_boundary_types(bt):
    from smaug_amd import configs
    rp = configs.stratified_params(ni=(8, 8, 8))
    rp.bcu = [bt, 4, 4]
    rp.bcl = [4, bt, 4]
    assert validate(rp) is None


def test_engine_refuses_to_run_without_library(monkeypatch, tmp_path):
    """No CPU fallback: a missing HIP library is an error, not a silent path."""
    monkeypatch.setattr(engine, "_lib", None)
    monkeypatch.setattr(engine, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(engine.SmaugError):
        engine.load()

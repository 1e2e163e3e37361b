"""Build recipe of the HIP engine: hipcc for gfx950 into smaug_amd/libsmaughip.so.

Run as ``python -m smaug_amd.build``.  The library is built in-tree so it
travels with the repository snapshot; nothing is installed.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libsmaughip.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("smaug_kernels.hip", "smaug_rhs.hip", "smaug_adj.hip", "smaug_host.hip", "smaug_io.cpp")]
ARCH = os.environ.get("SMAUG_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # fused multiply-add allowed: parity with the unfused oracle is held to a
    # tolerance (relative L2 per variable), not bitwise
    "-ffp-contract=fast",
    # no packed FP32 instructions (they are not used on purpose anywhere)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-pass-failed",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the engine)")


def build(verbose: bool = False) -> str:
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", LIB + ".tmp", *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed to build libsmaughip.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))

"""TEST INFRASTRUCTURE: compiles smaug_amd/csrc for the host against the HIP
emulation in this directory (g++), into tests/hostemu/_build/."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "smaug_amd", "csrc")
OUT = os.path.join(HERE, "_build", "libsmaughip_host.so")
SOURCES = ["smaug_kernels.hip", "smaug_rhs.hip", "smaug_adj.hip", "smaug_host.hip"]


def build() -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    srcs = [os.path.join(CSRC, f) for f in SOURCES]
    deps = srcs + [os.path.join(CSRC, f) for f in os.listdir(CSRC)] + [
        os.path.join(HERE, "emu.cpp"), os.path.join(HERE, "hip", "hip_runtime.h")]
    if os.path.exists(OUT) and os.path.getmtime(OUT) >= max(os.path.getmtime(d) for d in deps):
        return OUT
    tmp = f"{OUT}.{os.getpid()}.tmp"  # parallel test workers may build at once
    cmd = ["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-I", HERE, "-o", tmp,
           "-x", "c++", os.path.join(HERE, "emu.cpp"), *srcs, os.path.join(CSRC, "smaug_io.cpp")]
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())

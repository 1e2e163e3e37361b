"""Register budget of the stage kernel (CPU: compiles smaug_rhs.hip for
gfx950 and reads the compiler's resource report).  The FP64 k_rhs is tuned
for three wavefronts per SIMD with no scratch: 169 VGPRs already halve the
number of resident tiles (measured: 6.6 -> 8.6 ms per launch at 256^3), so a
change that crosses 168 must show up here rather than in the benchmark.  The
FP32 (mixed-precision) k_rhs runs five wavefronts per SIMD with at most a
small spill (measured: 4.0 ms against 4.3 ms at four wavefronts without a
spill and 4.6 ms at six with a 72-byte spill)."""
from __future__ import annotations

import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def resources(src_name, tmp_path):
    from smaug_amd.build import HIPCC_FLAGS

    flags = [f for f in HIPCC_FLAGS if f not in ("-shared", "-fPIC")]
    src = os.path.join(ROOT, "smaug_amd", "csrc", src_name)
    p = subprocess.run([HIPCC, *flags, "--cuda-device-only", "-c", src, "-o", str(tmp_path / "k.o"),
                        "-Rpass-analysis=kernel-resource-usage"], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-3000:]
    report = {}
    name = None
    for line in p.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            report[name] = {}
            continue
        m = re.search(r"remark:\s+(VGPRs|AGPRs|ScratchSize \[bytes/lane\]|Occupancy \[waves/SIMD\]): (\d+)", line)
        if m and name:
            report[name][m.group(1).split()[0]] = int(m.group(2))
    return report


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="needs hipcc")
def test_k_rhs_occupancy_and_scratch(tmp_path):
    from smaug_amd.build import ARCH

    report = resources("smaug_rhs.hip", tmp_path)
    k64 = {k: v for k, v in report.items() if "k_rhs_tiledILi3EdE" in k}
    k32 = {k: v for k, v in report.items() if "k_rhs_tiledILi3EfE" in k}
    assert len(k64) == 1 and len(k32) == 1, report
    for k, v in k64.items():
        assert v["Occupancy"] >= 3, (k, v)
        assert v["ScratchSize"] == 0, (k, v)
    for k, v in k32.items():
        assert v["Occupancy"] >= 5, (k, v)
        assert v["ScratchSize"] <= 16, (k, v)
    # the tangent-linear (Dual) instantiations: no scratch; the 3D one's
    # 105 KB of LDS allow one workgroup per CU, so its registers may run into
    # the accumulation file (256 + 10 at gfx950); the mixed-precision ones
    # (DualF) keep two wavefronts per SIMD or more
    kd = {k: v for k, v in report.items() if "k_rhs_tiled" in k and "4DualE" in k}
    kf = {k: v for k, v in report.items() if "k_rhs_tiled" in k and "5DualFE" in k}
    assert len(kd) == 2 and len(kf) == 2, report
    for k, v in kd.items():
        assert v["VGPRs"] + v.get("AGPRs", 0) <= 512 and v["ScratchSize"] == 0, (k, v)
    for k, v in kf.items():
        assert v["Occupancy"] >= 2 and v.get("AGPRs", 0) == 0 and v["ScratchSize"] == 0, (k, v)
    assert ARCH == "gfx950"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="needs hipcc")
def test_adjoint_kernels_plain_registers(tmp_path):
    """The adjoint kernels use no accumulation registers and no scratch and
    keep at least three wavefronts per SIMD (each term group is its own
    launch so no instantiation needs the register file of a whole step)."""
    report = resources("smaug_adj.hip", tmp_path)
    assert len(report) > 10
    for k, v in report.items():
        assert v.get("AGPRs", 0) == 0, (k, v)
        assert v["ScratchSize"] == 0, (k, v)
        assert v["Occupancy"] >= 3, (k, v)

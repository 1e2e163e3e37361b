"""CPU checks of the HIP kernels compiled for the host (tests/hostemu): the
same C ABI and kernel sources as the GPU library, run with every workgroup
thread as a fiber and guard pages around every buffer, so

  * an out-of-bounds access of any kernel ends the check (run twice: buffers
    placed against their trailing and their leading guard page);
  * the forward step matches the oracle (relative L2 <= 1e-12 per variable),
    and the mixed-precision step stays within 1e-5 of it;
  * the adjoint step (smaug_adjoint_step) matches Richardson-extrapolated
    central differences (h = 1e-6 relative) of the oracle with its
    hyperdiffusion coefficients replayed, to 1e-6 relative, per evolved
    field and for the driver amplitude map, for every term group;
  * the torch.autograd.Function path (smaug_amd.adjoint) gives the same
    gradient for every checkpoint interval and matches the differences;
  * the tangent-linear run (smaug_tl_*) matches the differences and is the
    transpose of the adjoint (dot products equal to roundoff).
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys

import pytest

from smaug_amd import params as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL = 1e-6

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


@pytest.fixture(scope="module")
def hostlib():
    sys.path.insert(0, ROOT)
    from tests.hostemu.build import build

    return build()


def run(hostlib, *args, align="end"):
    env = dict(os.environ, SMAUG_HIP_LIB=hostlib, SMAUG_ADJOINT="1", EMU_ALIGN=align, OMP_NUM_THREADS="1")
    print(f"[hostemu] {' '.join(args)} align={align}", flush=True)
    p = subprocess.run([sys.executable, "-m", "tests.hostemu.check", *args], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, f"host check died ({p.returncode}):\n{p.stdout[-2000:]}\n{p.stderr[-4000:]}"
    return json.loads(p.stdout.strip().splitlines()[-1])


def close(pair):
    got, want = pair
    return abs(got - want) <= TOL * abs(want)


@pytest.mark.parametrize("case", ["s2r", "s2e", "s3r", "s3m", "ot", "s2d"])
@pytest.mark.parametrize("align", ["end", "start"])
def test_forward_matches_oracle(hostlib, case, align):
    r = run(hostlib, "fwd", case, align=align)
    assert max(r["rel_l2"]) <= 1e-12, r


@pytest.mark.parametrize("kind", ["nan", "negative_rho_t"])
def test_blowup_is_reported_and_cleared(hostlib, kind):
    """k_prim's non-finite wave-speed flag reaches the caller as
    SMAUG_E_NUMERIC (-4) on the next synchronising call, and a state set
    afterwards clears it."""
    r = run(hostlib, "blowup", kind)
    assert r["raised"] is not None and "(-4)" in r["raised"], r
    assert r["after_reset"] == "ok", r


@pytest.mark.parametrize("case", ["s3e", "s3r", "s2r", "ot"])
def test_mixed_precision_matches_oracle(hostlib, case):
    """The FP32 stencil path (shadows, LDS boxes and arithmetic) stays finite
    and within 1e-5 relative L2 per variable of the FP64 oracle."""
    r = run(hostlib, "mixed", case)
    assert r["finite"], r
    assert max(r["rel_l2"]) <= 1e-5, r


@pytest.mark.parametrize("case", ["s2r", "s2e", "s3r", "ot"])
@pytest.mark.parametrize("align", ["end", "start"])
def test_adjoint_matches_replayed_differences(hostlib, case, align):
    r = run(hostlib, "adj", case, align=align)
    for f, pair in enumerate(r["fields"]):
        assert close(pair), (f, pair)
    if "amap" in r:
        assert close(r["amap"]), r["amap"]


def transposed(pair, tol=1e-10):
    """The tangent-linear derivative and the adjoint's dot product in one
    direction: the same linear map read forwards and transposed, so equal up
    to the roundoff of the two sums."""
    got, want = pair
    return abs(got - want) <= tol * abs(want)


@pytest.mark.parametrize("case", ["s2r", "s2e", "s3r", "ot", "tor", "spi", "bc:2d_fixed_asymm:0"])
def test_tangent_linear_matches_differences_and_adjoint(hostlib, case):
    """The tangent-linear run (smaug_tl_*): its values reproduce the oracle's
    loss, its directional derivatives of the line loss match the replayed
    differences of the oracle (1e-6) and equal the adjoint's dot products in
    the same directions (1e-10), per evolved field and for the driver map; in
    a direction of the background fields and along the background's controls
    (photospheric temperature, transition height: the builder's tangent
    against the oracle rerun on the rebuilt state) it matches the
    differences too, and for the cases without a map in the scalar driver
    amplitude; the engine's own state is left as it was."""
    r = run(hostlib, "tl", case)
    assert abs(r["loss"][0] - r["loss"][1]) <= 1e-12 * abs(r["loss"][1]), r["loss"]
    for f, pair in enumerate(r["fields"]):
        assert close(pair), (f, pair)
    for f, pair in enumerate(r["adjoint"]):
        assert transposed(pair), (f, pair)
    if "amap" in r:
        assert close(r["amap"]), r["amap"]
        assert transposed(r["amap_adjoint"]), r["amap_adjoint"]
    if r["bg"][1] == 0.0:  # no background (Orszag-Tang)
        assert r["bg"][0] == 0.0, r["bg"]
    else:
        assert close(r["bg"]), r["bg"]
    for name, pair in r.get("controls", {}).items():  # photospheric T, transition height
        assert close(pair), (name, pair)
    if "amplitude" in r:  # the scalar driver amplitude (cases without a map)
        assert close(r["amplitude"]), r["amplitude"]
    assert r["state_untouched"]


def test_tangent_linear_refusals_and_precision(hostlib):
    """The tangent-linear run refuses the divB source terms (-5), a step
    before its start, a driver-amplitude direction before its start, not
    finite, with a driver map or without a driver, and a driver-map
    direction without a map (-1); from a
    mixed-precision engine (FP32 pairs in the stencils) its values and
    tangents stay within 1e-5 of the FP64 run's after 10 steps, and are not
    identical to them."""
    r = run(hostlib, "tl_refusals")
    assert "failed (-5)" in r["divb"], r
    assert "failed (-1)" in r["step_before_begin"], r
    for key in ("amplitude_before_begin", "amplitude_not_finite", "amplitude_with_map", "amplitude_without_driver"):
        assert "failed (-1)" in r[key], (key, r)
    assert "failed (-1)" in r["map_direction_without_map"], r
    for key in ("mixed_values_rel_l2", "mixed_tangents_rel_l2"):
        assert max(r[key]) <= 1e-5 and max(r[key]) > 0, (key, r[key])


def test_tangent_torch_forward_mode(hostlib):
    """The forward-mode torch binding (smaug_amd.tangent.state_loss) on CPU
    tensors: the loss equals the oracle's; one dual run in a direction of
    the state, the background and the driver amplitude together matches the
    sum of the replayed differences (1e-6); an amplitude other than the
    engine's is refused."""
    r = run(hostlib, "tl_torch")
    assert abs(r["loss"][0] - r["loss"][1]) <= 1e-12 * abs(r["loss"][1]), r["loss"]
    assert close(r["all"]), r["all"]
    assert "driver amplitude" in r["wrong_amplitude"], r


BC_CASES = ["bc:2d_symm_cont:1", "bc:2d_fixed_asymm:0", "bc:3d_cont_symm_fixed:1", "bc:3d_asymm_fixed_cont:1"]


@pytest.mark.parametrize("case", BC_CASES)
def test_boundary_types_forward_and_adjoint(hostlib, case):
    """cont/fixed/symm/asymm tables: forward parity, and the adjoint's fold
    (signs of asymm images, no cotangent through fixed ghosts)."""
    r = run(hostlib, "fwd", case)
    assert max(r["rel_l2"]) <= 1e-12, r
    r = run(hostlib, "adj", case)
    for f, pair in enumerate(r["fields"]):
        assert close(pair), (f, pair)
    assert close(r["amap"]), r["amap"]


@pytest.mark.parametrize("group", ["flux", "eterm", "grav", "hrho", "he", "hmom", "hb"])
@pytest.mark.parametrize("case", ["s2r", "s3r"])
def test_adjoint_term_groups(hostlib, case, group):
    t = {"flux": P.T_FLUX, "eterm": P.T_ETERM, "grav": P.T_GRAV, "hrho": P.T_HRHO, "he": P.T_HE,
         "hmom": P.T_HMOM, "hb": P.T_HB}[group]
    r = run(hostlib, "adj", case, hex(t | P.T_SOURCE))
    for f, pair in enumerate(r["fields"]):
        if pair[1] == 0.0:
            assert pair[0] == 0.0 or abs(pair[0]) < 1e-12, (f, pair)
            continue
        assert close(pair), (f, pair)
    assert close(r["amap"]), r["amap"]


@pytest.mark.parametrize("rkon", [1, 0])
def test_autograd_function_and_checkpointing(hostlib, rkon):
    pytest.importorskip("torch")
    r = run(hostlib, "autograd", str(rkon))
    assert abs(r["loss"] - r["oracle_loss"]) <= 1e-12 * abs(r["oracle_loss"])
    assert r["every_identical"]
    assert close(r["w0"]), r["w0"]
    assert close(r["amap"]), r["amap"]


def test_short_caller_buffers_are_refused(hostlib):
    """Every entry point that takes a caller's device buffer checks that the
    buffer holds what the call reads or writes (hipMemGetAddressRange) and
    returns SMAUG_E_ARG otherwise."""
    r = run(hostlib, "ranges")
    assert set(r) == {"save_pert", "load_pert", "step_taped", "adjoint_step", "set_driver_map", "cfl_noalloc"}
    for name, msg in r.items():
        assert "failed (-1)" in msg and "past the end of its allocation" in msg, (name, msg)


@pytest.mark.parametrize("case", ["tor", "spi", "spe"])
def test_reference_drivers_forward_and_adjoint(hostlib, case):
    """fluxtube1's torsional driver (3D RK4) and spicule1's (2D RK4 and
    Euler, over the step whose counter is 1): forward parity with the oracle
    and the adjoint against the replayed differences."""
    r = run(hostlib, "fwd", case)
    assert max(r["rel_l2"]) <= 1e-12, r
    r = run(hostlib, "adj", case)
    for f, pair in enumerate(r["fields"]):
        assert close(pair), (f, pair)

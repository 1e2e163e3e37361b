#!/usr/bin/env python3
"""Benchmark of the SMAUG MHD time step on one MI355X.

Workload: the 3D gravity-stratified solar atmosphere (hydrostatic tanh
temperature profile, uniform vertical field, contcd4 boundaries,
hyperdiffusion, vertical-velocity driver) at 256^3 interior cells (260^3
with ghosts), four-step RK, with a seeded smoothed perturbation so every
hyperdiffusion branch is live.  One "step" is one RK4 time step.

Prints one JSON line: whole-job cell-updates/s (interior cells x steps /
wall time, state resident in HBM), the roofline of the dominant kernel
(k_rhs) measured with HIP events on the engine's stream, and a CPU baseline
(the oracle, a restatement of the reference algorithm) timed on a bounded
sample on this host.
"""
from __future__ import annotations

import argparse
import json

import numpy as np
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("cell-updates/s (RK4) at 256³, mixed vs FP64 on GPU; adjoint/forward time ratio")
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per interior cell of one k_rhs launch (one RK stage,
# pass B of the two-pass stage): 1664 B per cell-update of an RK4 step minus
# the wave-speed pass (13 doubles per stage), divided over the 4 stages.
RHS_BYTES_PER_CELL = (1664 - 4 * 13 * 8) / 4.0
PRIM_BYTES_PER_CELL = (13 + 4) * 8.0  # k_prim (FP64): reads 13 doubles per cell, writes v and q_e
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_rhs_stratified256_traffic.txt")


def read_traffic():
    """HBM bytes per k_rhs dispatch from the committed rocprofv3 --pmc summary."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE, encoding="utf-8") as fh:
        for line in fh:
            m = re.match(r"\s*traffic_bytes_per_launch\s*[:=]\s*([0-9.eE+-]+)", line)
            if m:
                return float(m.group(1))
    return None


def cpu_baseline(size: int, steps: int):
    """Oracle (CPU restatement of the reference step) on a bounded sample."""
    from oracle.pyoracle import Oracle, num_threads
    from smaug_amd import configs

    rp = configs.stratified_params(ni=(size, size, size), rkon=1)
    w, pos = configs.stratified(rp, perturb=20261202)
    orc = Oracle(rp, w, pos)
    t0 = time.perf_counter()
    orc.step(steps)
    dt = time.perf_counter() - t0
    orc.close()
    return {
        "value": size ** 3 * steps / dt,
        "unit": "cell-updates/s",
        "cores": num_threads(),
        "kind": "port",
        "sample": f"oracle RK4, {size}^3 interior, {steps} step(s), same stratified workload, {dt:.1f} s",
    }


COPY_PROBE_BYTES = 256 << 20


def copy_probe(torch):
    """A plain copy of known size (reads and writes COPY_PROBE_BYTES each):
    under rocprofv3 --pmc it gives the scale of FETCH_SIZE / WRITE_SIZE."""
    src = torch.rand(COPY_PROBE_BYTES // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(f"[bench] copy probe: 3 x {COPY_PROBE_BYTES} B read + written", file=sys.stderr, flush=True)
    del src, dst


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=256, help="interior cells per dim")
    ap.add_argument("--kernel-steps", type=int, default=3, help="steps timed per kernel with events")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tl-steps", type=int, default=3, help="tangent-linear steps timed (0 skips)")
    ap.add_argument("--cpu-size", type=int, default=96)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--mixed-check-steps", type=int, default=2000,
                    help="also run FP64, mixed and FP64 from a one-ulp-perturbed state from the initial state "
                         "for this many steps (2000: the horizon of the mixed-precision configuration) and "
                         "report the per-variable differences; 0 skips it")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a plain 256 MiB -> 256 MiB elementwise copy first (HBM counter calibration)")
    args = ap.parse_args(argv)

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build holds the whole domain on a single MI355X "
              "(multi-GPU halo exchange is not part of it). Run with --gpus 1.", file=sys.stderr)
        return 2

    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU (torch.cuda.is_available() is False)", file=sys.stderr)
        return 1
    from smaug_amd import configs
    from smaug_amd.engine import Engine
    from smaug_amd.params import PREC_MIXED, field_names

    if args.copy_probe:
        copy_probe(torch)
    S = args.size
    rp = configs.stratified_params(ni=(S, S, S), rkon=1)
    w, pos = configs.stratified(rp, perturb=20261202)
    print(f"[bench] state built: {S}^3 interior, {rp.ncells} cells incl. ghosts", file=sys.stderr, flush=True)
    cells = S ** 3

    def timed(eng, label):
        eng.step(args.warmup)
        eng.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.step(args.steps)
        eng.sync()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        print(f"[bench] {label}: timed {args.steps} steps, {wall / args.steps * 1e3:.3f} ms/step",
              file=sys.stderr, flush=True)
        return wall

    # FP64 (the plain form): the headline value
    eng = Engine(rp, w, pos)
    wall = timed(eng, "fp64")
    value = cells * args.steps / wall
    ms_step = wall / args.steps * 1e3
    state64 = eng.state()  # before the event-timed steps, to compare with the mixed run
    ms_prim, ms_rhs, ms_evstep = eng.time_kernels(args.kernel_steps)
    # tangent-linear step (forward-mode derivative, dual-number stage kernel)
    # from the same state: one untimed step, then timed steps
    def time_tl(e, label):
        e.tl_begin(np.zeros((2 + 2 * rp.ndim, rp.ncells)))
        e.tl_step(1)
        e.sync()
        t0 = time.perf_counter()
        e.tl_step(args.tl_steps)
        e.sync()
        ms = (time.perf_counter() - t0) / args.tl_steps * 1e3
        e.tl_end()
        print(f"[bench] {label}: timed {args.tl_steps} steps, {ms:.3f} ms/step", file=sys.stderr, flush=True)
        return ms

    tl = None
    if args.tl_steps > 0:
        ms_tl = time_tl(eng, "tangent-linear")
        tl = {"ms_per_step": ms_tl, "over_forward": ms_tl / ms_step, "steps": args.tl_steps,
              "note": "FP64 tangent-linear step (values and tangents, frozen nu, wave-speed maxima and dt)"}
    eng.close()

    # mixed precision: FP32 stencil shadows, FP64 state/CFL/combine, same
    # initial state and step count; error against the FP64 run
    rpm = configs.stratified_params(ni=(S, S, S), rkon=1)
    rpm.precision = PREC_MIXED
    engm = Engine(rpm, w, pos)
    del w, pos
    wallm = timed(engm, "mixed")
    statem = engm.state()
    msm_prim, msm_rhs, _ = engm.time_kernels(args.kernel_steps)
    if tl is not None:  # the mixed-precision tangent-linear step (FP32 pairs in the stencils)
        ms_tlm = time_tl(engm, "tangent-linear mixed")
        tl["mixed_ms_per_step"] = ms_tlm
        tl["mixed_speedup_vs_fp64"] = tl["ms_per_step"] / ms_tlm
    engm.close()
    k = 2 + 2 * rp.ndim

    def diffs(a, b):
        num = np.sqrt(((a[:k] - b[:k]) ** 2).sum(axis=1))
        den = np.sqrt((b[:k] ** 2).sum(axis=1))
        mx = np.abs(a[:k] - b[:k]).max(axis=1) / np.maximum(np.abs(b[:k]).max(axis=1), 1e-300)
        return (num / np.where(den == 0, 1.0, den)).tolist(), mx.tolist()

    rel, mxe = diffs(statem, state64)
    del state64, statem
    mixed = {
        "value": cells * args.steps / wallm,
        "unit": "cell-updates/s",
        "ms_per_step": wallm / args.steps * 1e3,
        "speedup_vs_fp64": wall / wallm,
        "k_rhs_ms_per_launch": msm_rhs,
        "k_prim_ms_per_launch": msm_prim,
        "rel_l2_vs_fp64_per_variable": dict(zip(field_names(rp.ndim)[:k], rel)),
        "max_err_over_max_fp64_per_variable": dict(zip(field_names(rp.ndim)[:k], mxe)),
        "steps_compared": args.warmup + args.steps,
    }
    if args.mixed_check_steps > 0:
        n = args.mixed_check_steps
        w, pos = configs.stratified(rp, perturb=20261202)
        # third run: FP64 from the initial state with every evolved value
        # moved by one ulp -- the scheme's own sensitivity to roundoff, the
        # floor under any comparison of two operation orders after n steps
        w_ulp = w.copy()
        rng = np.random.default_rng(1)
        w_ulp[:k] = np.nextafter(w_ulp[:k], np.where(rng.random(w_ulp[:k].shape) < 0.5, -np.inf, np.inf))
        states = []
        for label, prec, w_init in (("fp64", 0, w), ("mixed", PREC_MIXED, w), ("fp64 one-ulp", 0, w_ulp)):
            rq = configs.stratified_params(ni=(S, S, S), rkon=1)
            rq.precision = prec
            e = Engine(rq, w_init, pos)
            for done in range(0, n, 250):
                e.step(min(250, n - done))
                e.sync()
                print(f"[bench] long run {label}: {min(done + 250, n)}/{n} steps",
                      file=sys.stderr, flush=True)
            states.append(e.state())
            e.close()
        del w, pos, w_ulp
        rel_n, mx_n = diffs(states[1], states[0])
        rel_u, mx_u = diffs(states[2], states[0])
        names = field_names(rp.ndim)[:k]
        mixed["long_run"] = {
            "steps": n,
            "rel_l2_vs_fp64_per_variable": dict(zip(names, rel_n)),
            "max_err_over_max_fp64_per_variable": dict(zip(names, mx_n)),
            "fp64_one_ulp_spread_rel_l2_per_variable": dict(zip(names, rel_u)),
            "fp64_one_ulp_spread_max_err_per_variable": dict(zip(names, mx_u)),
        }
        del states
    rhs_bytes = RHS_BYTES_PER_CELL * cells
    achieved = rhs_bytes / (ms_rhs * 1e-3) / 1e9
    roof = {
        "kernel": "k_rhs",
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": read_traffic(),
        "algorithmic_bytes_per_launch": rhs_bytes,
        "ms_per_launch": ms_rhs,
        "k_prim_ms_per_launch": ms_prim,
        "k_prim_GBps": PRIM_BYTES_PER_CELL * rp.ncells / (ms_prim * 1e-3) / 1e9,
        "event_ms_per_step": ms_evstep,
    }
    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_size, args.cpu_steps)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"3D gravity-stratified atmosphere {S}^3 interior ({S + 4}^3 with ghosts), "
                        "RK4, FP64, hyperdiffusion, vertical driver, seeded perturbation",
            "interior_cells": cells,
            "precision": "fp64",
            "scheme": "rk4",
        },
        "mixed": mixed,
        "tangent_linear": tl,
        "adjoint": {
            "adjoint_over_forward": None,
            "status": "not measured: the adjoint entry points are gated off on the GPU (SMAUG_ADJOINT) "
                      "until the GPU fault of their first MI355X run is explained; they are checked in "
                      "the host build (tests/test_hostemu.py), where their dot products equal the "
                      "tangent-linear run's to roundoff; forward-mode derivatives run on the GPU "
                      "(tangent_linear above)",
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())

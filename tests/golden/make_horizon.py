"""Writes the long-horizon fixtures of the parity tests (run from the repo
root: ``python tests/golden/make_horizon.py [name ...]``; a few minutes per
3D case on 8 cores, seconds per 2D case).

The stratified configurations are run for their own horizons by the CPU
oracle: the 3D one for its 1000 RK4 steps on a 32 x 24 x 24 interior grid,
the 2D slice (dim 0 vertical, twice the cells of dim 1) for the 1500 steps of
its loss on 64 x 32.  Each run is made twice: from the configuration's state
and from the same state with every nonzero input value moved by one ulp (up
or down at random).  The second run measures how far the scheme itself
carries a roundoff-level difference over the horizon -- the floor under any
comparison of two operation orders -- and is stored with the fixture.

A full 3D state is 1.8 MB, so a 3D fixture keeps a sample of cells: two
whole interior planes through the middle of the grid (dims 0-1 and dims 0-2,
including the cells next to every face they cross) and 1536 seeded random
interior cells; a 2D fixture keeps every interior cell.  Each fixture stores
the evolved fields of those cells after the horizon, the per-variable
one-ulp spread on the sample and on the whole grid, the clock, and a SHA-256
of the inputs.
"""
from __future__ import annotations

import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from smaug_amd import configs  # noqa: E402
from smaug_amd.params import npert  # noqa: E402

NI = (32, 24, 24)
NI2 = (64, 32)


def case_perturbed():
    """the stratified configuration with its driver and a seeded 1e-3
    perturbation of every evolved field"""
    rp = configs.stratified_params(ni=NI, rkon=1)
    w, pos = configs.stratified(rp, perturb=20261202, amp=1e-3)
    return rp, w, pos


def case_driven():
    """the stratified configuration as specified: zero initial perturbation,
    everything driven by the vertical-velocity driver"""
    rp = configs.stratified_params(ni=NI, rkon=1)
    w, pos = configs.stratified(rp)
    return rp, w, pos


def case_slice_driven():
    """the 2D stratified slice driven from rest (the Gaussian driver; the
    per-column amplitude map is a control of the differentiable path)"""
    rp = configs.stratified_params(ni=NI2, ndim=2, rkon=1)
    w, pos = configs.stratified(rp)
    return rp, w, pos


# name: (builder, steps)
CASES = {
    "strat3d_32x24x24_rk4_1000_perturbed": (case_perturbed, 1000),
    "strat3d_32x24x24_rk4_1000_driven": (case_driven, 1000),
    "strat2d_64x32_rk4_1500_driven": (case_slice_driven, 1500),
}


def input_digest(w: np.ndarray, pos: np.ndarray) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(w, dtype=np.float64).tobytes())
    h.update(np.ascontiguousarray(pos, dtype=np.float64).tobytes())
    return h.hexdigest()


def sample_cells(rp) -> np.ndarray:
    """flat indices (dim 0 fastest) of the sampled interior cells, sorted"""
    n = rp.n
    if rp.ndim == 2:
        j, i = np.meshgrid(np.arange(n[1]), np.arange(n[0]), indexing="ij")
        inner = (i >= 2) & (i < n[0] - 2) & (j >= 2) & (j < n[1] - 2)
        return np.flatnonzero(inner.ravel()).astype(np.int64)
    k, j, i = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    inner = (i >= 2) & (i < n[0] - 2) & (j >= 2) & (j < n[1] - 2) & (k >= 2) & (k < n[2] - 2)
    plane = inner & ((k == n[2] // 2) | (j == n[1] // 2))
    rest = np.flatnonzero(inner.ravel() & ~plane.ravel())
    pick = np.random.default_rng(20261203).choice(rest, 1536, replace=False)
    return np.sort(np.concatenate([np.flatnonzero(plane.ravel()), pick])).astype(np.int64)


def one_ulp(w: np.ndarray) -> np.ndarray:
    """every nonzero value moved one ulp up or down (seeded)"""
    wu = w.copy()
    nz = wu != 0
    up = np.random.default_rng(1).random(wu.shape) < 0.5
    wu[nz] = np.nextafter(wu[nz], np.where(up, np.inf, -np.inf)[nz])
    return wu


def rel_l2(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    num = np.sqrt(((a - b) ** 2).sum(axis=1))
    den = np.sqrt((b ** 2).sum(axis=1))
    return num / np.where(den == 0, 1.0, den)


def path(name: str) -> str:
    return os.path.join(HERE, name + ".npz")


def main(names):
    from oracle.pyoracle import Oracle

    for name in names:
        build, steps = CASES[name]
        rp, w, pos = build()
        k = npert(rp.ndim)
        idx = sample_cells(rp)
        t0 = time.time()
        out = {}
        for label, w_in in (("ref", w), ("ulp", one_ulp(w))):
            o = Oracle(rp, w_in, pos)
            o.step(steps)
            out[label] = o.state()[:k]
            if label == "ref":
                clock = np.array(o.clock())
            o.close()
            print(f"{name}: {label} run done, {time.time() - t0:.0f} s", flush=True)
        ref, ulp = out["ref"], out["ulp"]
        if not (np.isfinite(ref).all() and np.isfinite(ulp).all()):
            raise SystemExit(f"{name}: the oracle's state is not finite after {steps} steps")
        spread_sample = rel_l2(ulp[:, idx], ref[:, idx])
        spread_full = rel_l2(ulp, ref)
        np.savez_compressed(path(name), idx=idx, w=ref[:, idx], spread_sample=spread_sample,
                            spread_full=spread_full, clock=clock, steps=steps, ni=np.array(rp.ni),
                            digest=np.array(input_digest(w, pos)))
        print(f"{name}: {idx.size} cells, spread sample {spread_sample.max():.3e} full "
              f"{spread_full.max():.3e}, {os.path.getsize(path(name))} bytes", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or list(CASES))

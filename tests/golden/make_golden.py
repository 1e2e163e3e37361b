"""Writes the golden fixtures of the parity tests (run from the repo root:
``python tests/golden/make_golden.py``).

The reference holds no result vectors for this path, so the fixtures are
outputs of the CPU oracle (oracle/, pinned by tests/test_oracle_pins.py) on
small seeded cases.  Each fixture stores the evolved fields after N steps,
the clock, and a SHA-256 of the inputs, so a change of the input generators
is reported as such rather than as a parity failure.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from cases import ot_case, strat_case  # noqa: E402
from smaug_amd.params import npert  # noqa: E402

CASES = {
    # name: (builder, steps)
    "ot16_euler": (lambda: ot_case(ni=16, rkon=0), 3),
    "ot16_rk4": (lambda: ot_case(ni=16, rkon=1), 2),
    "strat2d_12x10_rk4": (lambda: strat_case(ni=(12, 10), ndim=2, rkon=1), 3),
    "strat3d_10x8x8_rk4": (lambda: strat_case(ni=(10, 8, 8), rkon=1), 2),
    "strat3d_10x8x8_euler": (lambda: strat_case(ni=(10, 8, 8), rkon=0), 2),
}


def input_digest(w: np.ndarray, pos: np.ndarray) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(w, dtype=np.float64).tobytes())
    h.update(np.ascontiguousarray(pos, dtype=np.float64).tobytes())
    return h.hexdigest()


def path(name: str) -> str:
    return os.path.join(HERE, name + ".npz")


def main():
    from oracle.pyoracle import Oracle

    for name, (build, steps) in CASES.items():
        rp, w, pos = build()
        o = Oracle(rp, w, pos)
        o.step(steps)
        out = o.state()[: npert(rp.ndim)]
        it, t, qt, dt = o.clock()
        o.close()
        np.savez_compressed(path(name), w=out, clock=np.array([it, t, qt, dt]), steps=steps,
                            digest=np.array(input_digest(w, pos)))
        print(f"{name}: {steps} steps, {out.shape}, {os.path.getsize(path(name))} bytes")


if __name__ == "__main__":
    main()

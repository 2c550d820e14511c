"""Writes the oracle's golden fixtures (regression pins of the CPU restatement).

The reference holds no golden vectors for this path (its tests cover geometry
primitives only), so these fixtures pin the oracle against its own past
output: a change that moves any bit of a neighbor set, density, position or
velocity fails tests/test_golden.py. What pins the oracle against the
reference's formulas is the closed-form and invariant suite
(tests/test_oracle_*.py).

Cases: jittered lattices 4^3, 8^3, 16^3 (spacing 0.005, jitter 0.1*spacing,
origin (0.3, 0.5, 0.3), seeds 41, 42, 43), zero initial velocity, reference
parameters (h 0.0125, 5 iterations, dt 1e-3, g -9.8); outputs after 1 and 3
steps and the first step's neighbor lists (full CSR for 4^3 and 8^3, per-row
count and id sum for 16^3).

Run from the repository root: python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import __graft_entry__ as G  # noqa: E402

CASES = {"lattice4": (4, 41), "lattice8": (8, 42), "lattice16": (16, 43)}


def make_case(side, seed):
    scene = G.load_package().scene
    O = G.load_oracle()
    p = O.make_params()
    x = scene.jittered_lattice((side, side, side), (0.3, 0.5, 0.3), 0.005, 0.1, seed=seed)
    v = np.zeros_like(x)
    off, el = O.neighbor_query(p, x)
    x1, v1, r1, _ = O.step(p, x, v, 1)
    x3, v3, r3, _ = O.step(p, x, v, 3)
    out = dict(x0=x, v0=v, x1=x1, v1=v1, rho1=r1, x3=x3, v3=v3, rho3=r3, seed=np.int64(seed),
               offsets=off.astype(np.int32))
    if side <= 8:
        rows = [np.sort(el[off[i]:off[i + 1]]) for i in range(len(x))]
        out["elems_sorted"] = np.concatenate(rows).astype(np.int32)
    else:
        out["row_id_sum"] = np.add.reduceat(el.astype(np.int64), off[:-1].astype(np.int64))
    return out


def main():
    here = Path(__file__).resolve().parent
    for name, (side, seed) in CASES.items():
        np.savez_compressed(here / f"{name}.npz", **make_case(side, seed))
        print("wrote", name)


if __name__ == "__main__":
    main()

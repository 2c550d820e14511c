"""The oracle reproduces its committed golden fixtures bit for bit
(tests/golden/make_golden.py describes the cases)."""
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = ["lattice4", "lattice8", "lattice16"]


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_golden(oracle, name):
    g = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    p = oracle.make_params()
    off, el = oracle.neighbor_query(p, g["x0"])
    np.testing.assert_array_equal(off, g["offsets"])
    if "elems_sorted" in g:
        rows = [np.sort(el[off[i]:off[i + 1]]) for i in range(len(off) - 1)]
        np.testing.assert_array_equal(np.concatenate(rows), g["elems_sorted"])
    else:
        np.testing.assert_array_equal(np.add.reduceat(el.astype(np.int64), off[:-1].astype(np.int64)),
                                      g["row_id_sum"])
    for steps in (1, 3):
        x, v, rho, _ = oracle.step(p, g["x0"], g["v0"], steps)
        np.testing.assert_array_equal(x, g[f"x{steps}"])
        np.testing.assert_array_equal(v, g[f"v{steps}"])
        np.testing.assert_array_equal(rho, g[f"rho{steps}"])

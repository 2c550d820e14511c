"""Helpers shared by the GPU parity tests: tolerances, tensor transfer,
order-independent neighbor-set comparison and the step-parity check."""
import sys

import numpy as np
import torch

RHO_REL = 1e-5
X_ABS = 1e-6
V_ABS = 1e-3


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to("cuda:0")


def row_keys(off, el):
    n = len(off) - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(off).astype(np.int64))
    return np.sort(rows * n + el.astype(np.int64))


def row_checksums(off, el):
    """Order-independent per-row (count, sum, xor) of SplitMix-mixed ids."""
    z = el.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z = z ^ (z >> np.uint64(31))
    cnt = np.diff(off)
    starts = off[:-1].astype(np.int64)
    nz = cnt > 0
    s = np.zeros(len(cnt), np.uint64)
    x = np.zeros(len(cnt), np.uint64)
    s[nz] = np.add.reduceat(z, starts[nz])
    x[nz] = np.bitwise_xor.reduceat(z, starts[nz])
    return cnt, s, x


def assert_same_sets(gpu_nl, off, el):
    """gpu_nl: a NeighborList (n row starts + elements); off: the oracle's
    n + 1 offsets."""
    assert gpu_nl.index.numel() == len(off) - 1 and gpu_nl.elements.numel() == off[-1]
    go = gpu_nl.offsets().cpu().numpy()
    ge = gpu_nl.elements.cpu().numpy()
    np.testing.assert_array_equal(go, off)
    np.testing.assert_array_equal(row_keys(go, ge), row_keys(off, el))


def rel_err(a, b):
    """|a - b| / |b| per element; 0 where equal (also 0 == 0), |a - b| where b == 0."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(a - b)
    den = np.where(np.abs(b) > 0, np.abs(b), 1.0)
    return np.where(d == 0, 0.0, d / den)


def assert_step_parity(gpu, f32, f64=None):
    """gpu, f32, f64: (x, v, rho) of the GPU, the FP32 restatement and the FP64
    instantiation on the same input.

    Without f64: x within X_ABS, v within V_ABS and rho within RHO_REL of the
    FP32 restatement. With f64, each quantity is held against FP64 to its bar
    or twice the FP32 restatement's own distance from FP64, whichever is
    larger: where hundreds of neighbor terms are summed (dense blobs) an fp32
    sum in another order is as far from the FP32 restatement as that is from
    FP64, and neither is closer to the exact result."""
    x, v, rho = gpu
    xo, vo, ro = f32[:3]
    ex32 = np.abs(x.astype(np.float64) - xo).max()
    ev32 = np.abs(v.astype(np.float64) - vo).max()
    if f64 is None:
        assert rel_err(rho, ro).max() <= RHO_REL, f"rho rel err {rel_err(rho, ro).max():.3e}"
        assert ex32 <= X_ABS, f"x err {ex32:.3e}"
        assert ev32 <= V_ABS, f"v err {ev32:.3e}"
        return
    x64, v64, r64 = (np.asarray(a, np.float64) for a in f64[:3])
    checks = (
        ("rho", rel_err(rho, r64).max(), rel_err(ro, r64).max(), RHO_REL),
        ("x", np.abs(x - x64).max(), np.abs(xo - x64).max(), X_ABS),
        ("v", np.abs(v - v64).max(), np.abs(vo - v64).max(), V_ABS),
    )
    say("; ".join(f"{name}: gpu-vs-fp64 {er:.3e}, fp32-oracle-vs-fp64 {floor:.3e}, bar {max(b, 2 * floor):.3e}"
                  for name, er, floor, b in checks)
        + f"; gpu-vs-fp32: rho {rel_err(rho, ro).max():.3e} x {ex32:.3e} v {ev32:.3e}")
    for name, er, floor, b in checks:
        assert er <= max(b, 2 * floor), f"{name} err {er:.3e} > {max(b, 2 * floor):.3e}"


def oracle_both(oracle, x, v, nsteps, **kw):
    f32 = oracle.step(oracle.make_params(**kw), x, v, nsteps)
    f64 = oracle.step(oracle.make_params(double=True, **kw), np.asarray(x, np.float64), np.asarray(v, np.float64),
                      nsteps)
    return f32, f64

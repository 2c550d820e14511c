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
    instantiation on the same input."""
    x, v, rho = gpu
    xo, vo, ro = f32[:3]
    ex = np.abs(x.astype(np.float64) - xo).max()
    ev = np.abs(v.astype(np.float64) - vo).max()
    if f64 is None:
        er, bar = rel_err(rho, ro).max(), RHO_REL
    else:
        floor = rel_err(ro, f64[2]).max()
        er, bar = rel_err(rho, f64[2]).max(), max(RHO_REL, 2 * floor)
        say(f"rho: gpu-vs-fp64 {er:.3e}, fp32-oracle-vs-fp64 {floor:.3e}, bar {bar:.3e}; "
            f"gpu-vs-fp32: rho {rel_err(rho, ro).max():.3e} x {ex:.3e} v {ev:.3e}")
    assert er <= bar, f"rho rel err {er:.3e} > {bar:.3e}"
    assert ex <= X_ABS, f"x err {ex:.3e}"
    assert ev <= V_ABS, f"v err {ev:.3e}"


def oracle_both(oracle, x, v, nsteps, **kw):
    f32 = oracle.step(oracle.make_params(**kw), x, v, nsteps)
    f64 = oracle.step(oracle.make_params(double=True, **kw), np.asarray(x, np.float64), np.asarray(v, np.float64),
                      nsteps)
    return f32, f64

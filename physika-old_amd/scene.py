"""Seeded particle-scene generators for the PBD fluid step.

The reference fills scenes with ``ParticleSystem::loadParticles(lo, hi, d)``,
a plain lattice loop (Dynamics/ParticleSystem/ParticleSystem.cpp:72-98). The
workloads here are jittered lattices, Gaussian blobs and batches of lattice
blocks, drawn from a counter-based SplitMix64 stream (Steele, Lea and Flood,
"Fast splittable pseudorandom number generators", OOPSLA 2014): output k of
seed s is mix(s + (k+1)*0x9E3779B97F4A7C15), so every value is a pure function
of (seed, index) and the arrays are identical on every machine.

Lattice order is x fastest: index = i + nx*(j + ny*k).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Outputs offset .. offset+count-1 of the SplitMix64 stream of ``seed``."""
    k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform01(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """float64 in [0, 1) with 53 random bits."""
    return (splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Standard normals by Box-Muller, two stream outputs per value."""
    u = uniform01(seed, 2 * count, 2 * offset).reshape(count, 2)
    u1 = 1.0 - u[:, 0]  # (0, 1]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u[:, 1])


def jittered_lattice(dims, origin, spacing, jitter_frac=0.1, seed=1, count=None) -> np.ndarray:
    """(n, 3) float32 lattice origin + spacing*(i, j, k), x fastest, with each
    coordinate displaced by U(-jitter_frac*spacing, +jitter_frac*spacing).
    ``count`` keeps only the first ``count`` lattice points."""
    nx, ny, nz = (int(d) for d in dims)
    n = nx * ny * nz if count is None else int(count)
    idx = np.arange(n, dtype=np.int64)
    ijk = np.stack([idx % nx, (idx // nx) % ny, idx // (nx * ny)], axis=1).astype(np.float64)
    jit = (2.0 * uniform01(seed, 3 * n).reshape(n, 3) - 1.0) * (jitter_frac * spacing)
    pos = np.asarray(origin, dtype=np.float64)[None, :] + spacing * ijk + jit
    return np.ascontiguousarray(pos.astype(np.float32))


def gaussian_blobs(nblobs, per_blob, sigma, seed, centre_lo=0.25, centre_hi=0.75, box=(0.0, 1.0)) -> np.ndarray:
    """nblobs*per_blob points: blob centres U([centre_lo, centre_hi]^3), points
    N(centre, sigma^2 I); a point outside ``box`` is redrawn from the stream."""
    c = centre_lo + (centre_hi - centre_lo) * uniform01(seed, 3 * nblobs).reshape(nblobs, 3)
    out = np.empty((nblobs * per_blob, 3), dtype=np.float64)
    lo, hi = box
    offset = 0
    for b in range(nblobs):
        need = per_blob
        got = []
        while need > 0:
            m = max(2 * need, 1024)
            z = normal(seed + 1 + b, 3 * m, offset).reshape(m, 3)
            offset += 3 * m
            p = c[b][None, :] + sigma * z
            ok = np.all((p >= lo) & (p <= hi), axis=1)
            p = p[ok][:need]
            got.append(p)
            need -= len(p)
        out[b * per_blob:(b + 1) * per_blob] = np.concatenate(got)
    return np.ascontiguousarray(out.astype(np.float32))


@dataclass
class BatchedScenes:
    counts: np.ndarray                 # (S,) particles per scene
    offsets: np.ndarray                # (S+1,) exclusive offsets into pos
    pos: np.ndarray                    # (sum, 3) float32, scenes concatenated
    spacing: float = 0.005
    meta: dict = field(default_factory=dict)


def batched_lattice_scenes(nscenes=256, nmin=2048, nmax=65536, seed=4, origin=(0.3, 0.5, 0.3),
                           spacing=0.005) -> BatchedScenes:
    """Scenes with n_s = floor(exp(U(ln nmin, ln nmax))) particles, each the
    first n_s points (x fastest) of a jittered cubic lattice block of side
    ceil(n_s^(1/3)) at ``origin`` in its own [0,1]^3, seeded seed + scene."""
    u = uniform01(seed, nscenes)
    counts = np.floor(np.exp(math.log(nmin) + u * (math.log(nmax) - math.log(nmin)))).astype(np.int64)
    parts = []
    for s, n in enumerate(counts):
        side = int(math.ceil(round(n ** (1.0 / 3.0), 9)))
        while side ** 3 < n:
            side += 1
        parts.append(jittered_lattice((side, side, side), origin, spacing, 0.1, seed + 1 + s, count=int(n)))
    offsets = np.zeros(nscenes + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(counts)
    return BatchedScenes(counts=counts, offsets=offsets, pos=np.concatenate(parts), spacing=spacing)


# Named workloads (spacing 0.005, h = 2.5*spacing, rest density 1000, dt 1e-3,
# 5 PBD iterations, gravity -9.8 on y, scene box [0,1]^3).
def dam_break_32k() -> np.ndarray:
    """32x32x32 = 32,768-particle jittered lattice at (0.10, 0.50, 0.10), seed 1."""
    return jittered_lattice((32, 32, 32), (0.10, 0.50, 0.10), 0.005, 0.1, seed=1)


def lattice_1m() -> np.ndarray:
    """128(x) x 64(y) x 128(z) = 1,048,576-particle jittered lattice at
    (0.18, 0.40, 0.18), seed 2."""
    return jittered_lattice((128, 64, 128), (0.18, 0.40, 0.18), 0.005, 0.1, seed=2)


def blobs_1m() -> np.ndarray:
    """8 Gaussian blobs x 131,072 = 1,048,576 particles, sigma 0.05, seed 3."""
    return gaussian_blobs(8, 131072, 0.05, seed=3)


def block_64k() -> np.ndarray:
    """64x32x32 = 65,536-particle jittered block at (0.3, 0.5, 0.3), seed 5."""
    return jittered_lattice((64, 32, 32), (0.3, 0.5, 0.3), 0.005, 0.1, seed=5)

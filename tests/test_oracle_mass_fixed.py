"""Pins of the oracle's massInv variant of DensityPBD and of the fixed-capacity
neighbor rows (CPU only)."""
import math

import numpy as np


def block(scene, dims=(8, 7, 6), spacing=0.0042, seed=5):
    return scene.jittered_lattice(dims, (0.3, 0.5, 0.3), spacing, 0.1, seed=seed)


def test_unit_inverse_masses_reproduce_the_plain_path(oracle, scene):
    x = block(scene)
    v = np.zeros_like(x)
    P = oracle.make_params(dx_mode=2)  # the literal scatter, as the massInv path uses
    off, el = oracle.neighbor_query(P, x)
    xa, va, ra = oracle.pbd_constrain(P, x, v, off, el)
    xb, vb, rb = oracle.pbd_constrain_minv(P, x, v, off, el, np.ones(len(x), np.float32))
    np.testing.assert_array_equal(xa, xb)
    np.testing.assert_array_equal(va, vb)
    np.testing.assert_array_equal(ra, rb)


def test_mass_weighted_momentum_is_conserved(oracle, scene):
    x = block(scene).astype(np.float64)
    n = len(x)
    v = np.zeros_like(x)
    minv = 0.5 + 1.5 * scene.uniform01(9, n)
    P = oracle.make_params(double=True)
    off, el = oracle.neighbor_query(oracle.make_params(), x.astype(np.float32))
    xo, vo, _ = oracle.pbd_constrain_minv(P, x, v, off, el, minv)
    assert np.abs(xo - x).max() > 1e-5  # the block is compressed: particles move
    m = 1.0 / minv
    p = (m[:, None] * (vo - v)).sum(0)
    scale = (m[:, None] * np.abs(vo - v)).sum(0)
    assert np.all(np.abs(p) <= 1e-10 * scale)


def test_two_particles_with_masses_closed_form(oracle):
    h = 0.0125
    r = 0.3 * h
    x = np.array([[0.5, 0.5, 0.5], [0.5 + r, 0.5, 0.5]])
    minv = np.array([2.0, 0.5])
    P = oracle.make_params(double=True, lam_rest=1.0, iters=1)
    off, el = oracle.neighbor_query(oracle.make_params(), x.astype(np.float32))
    xo, vo, rho = oracle.pbd_constrain_minv(P, x, np.zeros_like(x), off, el, minv)
    # closed form in FP64
    d = 0.005
    mass = oracle.lib().orc_scaling_factor_d(d, h) * 1000.0 * d ** 3
    w = lambda q: 15.0 / (math.pi * h ** 3) * (1 - q) ** 3  # noqa: E731
    grad = -45.0 / (math.pi * h ** 3) * (1 - r / h) ** 2
    rho_c = mass * (w(0.0) + w(r / h))
    np.testing.assert_allclose(rho, [rho_c, rho_c], rtol=1e-12)
    g2 = grad ** 2  # |g|^2 for the pair (unit direction)
    lam = np.array([-(rho_c - 1.0) / (g2 * minv[1] + g2 * minv[0] + 0.1),
                    -(rho_c - 1.0) / (g2 * minv[0] + g2 * minv[1] + 0.1)])
    lam = np.minimum(lam, 0.0)
    assert lam.max() < 0  # lambda active
    dp = 10.0 * (-r) * (lam[0] + lam[1]) * grad / r  # x-component of dp_01 (x_0 - x_1 = -r)
    np.testing.assert_allclose(xo[0, 0] - x[0, 0], 2 * dp * minv[0], rtol=1e-10)
    np.testing.assert_allclose(xo[1, 0] - x[1, 0], -2 * dp * minv[1], rtol=1e-10)


def test_fixed_rows_hold_the_nearest(oracle, scene):
    x = block(scene, (9, 9, 9), 0.005, 11)
    P = oracle.make_params()
    off, el = oracle.neighbor_query(P, x)
    for limit in (1, 5, 16, 200):
        counts, elems = oracle.neighbor_query_fixed(P, x, limit)
        sizes = np.diff(off)
        np.testing.assert_array_equal(counts, np.minimum(sizes, limit))
        for i in range(0, len(x), 37):
            row = el[off[i]:off[i + 1]]
            d = np.sqrt(((x[i].astype(np.float64) - x[row].astype(np.float64)) ** 2).sum(1))
            order = np.argsort(d, kind="stable")
            k = counts[i]
            if k < len(row) and np.isclose(d[order[k - 1]], d[order[k]], rtol=1e-6):
                continue  # a tie at the boundary: either is a valid choice
            assert set(elems[i, :k]) == set(row[order[:k]])

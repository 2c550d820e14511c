"""Pins of the oracle's ImplicitViscosity and box boundary (CPU only).

Viscosity: the weight's closed form, invariance of a uniform velocity field,
an isolated particle, and the two-particle Jacobi recursion in closed form.
Box: the sampled field against the analytic distance to the box, the
interpolated distance and normal near each wall, the projection of a particle
outside a wall in closed form, the jitter's range, and the step's
composition with both modules."""
import math

import numpy as np
import pytest

H = 0.0125


def test_vis_weight_closed_form(oracle):
    c = 45.0 / (13.0 * math.pi * H ** 3)
    for q in (0.0, 0.25, 0.5, 0.999):
        r = np.float32(q * H)
        qf = float(np.float32(r) / np.float32(H))  # the fp32 ratio the reference forms
        assert oracle.vis_weight(r, H) == pytest.approx(c * (1 - qf), rel=2e-6)
    assert oracle.vis_weight(H, H) == 0.0
    assert oracle.vis_weight(1.0001 * H, H) == 0.0


def lattice_list(oracle, scene, dims=(6, 6, 6), seed=3):
    x = scene.jittered_lattice(dims, (0.3, 0.5, 0.3), 0.005, 0.1, seed=seed)
    off, el = oracle.neighbor_query(oracle.make_params(), x)
    return x, off, el


def test_viscosity_keeps_uniform_velocity(oracle, scene):
    x, off, el = lattice_list(oracle, scene)
    v = np.tile(np.array([[0.3, -1.2, 0.7]], np.float32), (len(x), 1))
    out = oracle.viscosity(oracle.make_params(), x, v, off, el)
    np.testing.assert_allclose(out, v, rtol=3e-6, atol=0)


def test_viscosity_isolated_particle_unchanged(oracle):
    x = np.array([[0.5, 0.5, 0.5]], np.float32)
    v = np.array([[1.0, 2.0, 3.0]], np.float32)
    off, el = oracle.neighbor_query(oracle.make_params(), x)
    out = oracle.viscosity(oracle.make_params(), x, v, off, el)
    np.testing.assert_array_equal(out, v)  # sum w = 0 -> b = 0: v/(1+0) + 0


def test_viscosity_two_particles_closed_form(oracle):
    r = 0.4 * H
    x = np.array([[0.5, 0.5, 0.5], [0.5 + r, 0.5, 0.5]], np.float32)
    v0 = np.array([[1.0, 0.0, -2.0], [-1.0, 0.5, 0.0]], np.float64)
    P = oracle.make_params()
    off, el = oracle.neighbor_query(P, x)
    out = oracle.viscosity(P, x, v0.astype(np.float32), off, el)
    b = 0.001 * 1.0 / H
    v = v0.copy()
    for _ in range(5):  # each particle's weighted mean is the other's velocity
        v = v0 / (1 + b) + v[::-1] * b / (1 + b)
    np.testing.assert_allclose(out, v, rtol=1e-6, atol=1e-7)


def analytic_box_distance(p, lo, hi):
    """Inside: distance to the nearest face (positive); outside: minus the
    Euclidean distance to the box."""
    p = np.asarray(p, np.float64)
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    inside = np.all((p > lo) & (p < hi), axis=1)
    d_in = np.minimum(p - lo, hi - p).min(axis=1)
    q = np.maximum(np.maximum(lo - p, p - hi), 0.0)
    d_out = -np.sqrt((q * q).sum(1))
    return np.where(inside, d_in, d_out)


def test_box_field_samples(oracle):
    lo, hi, sp = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.02
    dims, left, h, d = oracle.box_sdf(lo, hi, sp)
    assert list(dims) == [61, 61, 61]  # floor(1/0.02) + 10 + 1
    np.testing.assert_allclose(left, -0.1, rtol=1e-6)
    k, j, i = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    pts = np.stack([left[0] + i.astype(np.float32) * h[0], left[1] + j.astype(np.float32) * h[1],
                    left[2] + k.astype(np.float32) * h[2]], -1).reshape(-1, 3).astype(np.float32)
    want = analytic_box_distance(pts, lo, hi)
    got = d.reshape(-1).astype(np.float64)
    inside = want > 0
    np.testing.assert_allclose(got[inside], want[inside], rtol=0, atol=1e-6)
    # outside, the reference's edge distance sqrt(|e|^2 - (e.s)^2) cancels: its
    # fp32 error reaches sqrt(8 eps |e|^2) ~ 1e-4 for samples close to an edge line
    err = np.abs(got[~inside] - want[~inside])
    assert err.max() <= 1.2e-4 and np.median(err) <= 1e-6


def test_box_field_distance_and_normal_near_walls(oracle):
    lo, hi, sp = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.005
    pts, normals = [], []
    for a in range(3):
        for side, nsign in ((0.003, -1.0), (0.995, 1.0)):
            p = [0.4, 0.6, 0.45]
            p[a] = side
            pts.append(p)
            n = [0.0, 0.0, 0.0]
            n[a] = nsign
            normals.append(n)
    pts = np.array(pts, np.float32)
    d, nrm = oracle.sdf_distance(lo, hi, sp, pts)
    np.testing.assert_allclose(d, analytic_box_distance(pts, lo, hi), atol=2e-6)
    np.testing.assert_allclose(nrm, np.array(normals), atol=1e-6)


def test_box_constrain_projection_closed_form(oracle):
    P = oracle.make_params()
    x = np.array([[-0.001, 0.5, 0.5], [0.5, 0.5, 0.5]], np.float32)
    v = np.array([[-1.0, 0.5, 0.0], [0.3, 0.2, 0.1]], np.float32)
    xo, vo = oracle.box_constrain(P, x, v)
    jit = oracle.box_jitter(0)
    assert 0.0 < jit <= 1e-4
    # the wall x = 0: d = -0.001, outward normal (-1, 0, 0)
    assert xo[0, 0] == pytest.approx(jit, abs=2e-7)
    assert xo[0, 1] == x[0, 1] and xo[0, 2] == x[0, 2]
    # normal part reflected and scaled by 1 - 0.95, tangential part kept
    np.testing.assert_allclose(vo[0], [0.05, 0.5, 0.0], atol=1e-6)
    # a particle inside, away from the walls: untouched
    np.testing.assert_array_equal(xo[1], x[1])
    np.testing.assert_array_equal(vo[1], v[1])


def test_box_jitter_range_and_determinism(oracle):
    j = np.array([oracle.box_jitter(i) for i in range(4096)])
    assert j.min() > 0 and j.max() <= 1e-4
    assert len(np.unique(j)) > 4000
    assert abs(j.mean() - 5e-5) < 3e-6
    assert oracle.box_jitter(17) == oracle.box_jitter(17)


def test_step_composes_viscosity_and_box(oracle, scene):
    """flags 1: the step's positions unchanged, velocities = viscosity on the
    list built at x^n; flags 2: the box first, then the plain step."""
    x = scene.jittered_lattice((10, 8, 10), (-0.004, 0.02, 0.3), 0.005, 0.1, seed=9)
    v = np.zeros_like(x)
    v[:, 0] = -0.5
    P0 = oracle.make_params()
    off, el = oracle.neighbor_query(P0, x)
    x1, v1, _, _ = oracle.step(P0, x, v, 1)
    xv, vv, _, _ = oracle.step(oracle.make_params(flags=oracle.FLAG_VISCOSITY), x, v, 1)
    np.testing.assert_array_equal(xv, x1)
    np.testing.assert_array_equal(vv, oracle.viscosity(P0, x1, v1, off, el))
    xb, vb = oracle.box_constrain(P0, x, v)
    assert (xb != x).any()  # some particles start beyond the wall x = 0
    x2, v2, _, _ = oracle.step(P0, xb, vb, 1)
    xs, vs, _, _ = oracle.step(oracle.make_params(flags=oracle.FLAG_BOX), x, v, 1)
    np.testing.assert_array_equal(xs, x2)
    np.testing.assert_array_equal(vs, v2)

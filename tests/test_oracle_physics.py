"""Oracle pins for the Spiky kernel, SummationDensity, the lambda and
displacement passes of DensityPBD, the integrator and the step order.

Each check is a closed form, an invariant or an independent emulation chosen
so that a plausible mistake (dropped factor, wrong sign, wrong order) fails."""
import math

import numpy as np
import pytest

H = 0.0125
D = 0.005


def w64(r, h=H):
    q = r / h
    return 0.0 if q > 1 else 15.0 / (math.pi * h ** 3) * (1 - q) ** 3


def dw64(r, h=H):
    q = r / h
    return 0.0 if q > 1 else -45.0 / (math.pi * h ** 3) * (1 - q) ** 2


# ---------------------------------------------------------------- Spiky kernel
@pytest.mark.parametrize("q", [0.0, 0.25, 0.5, 0.9, float(np.nextafter(np.float32(1), np.float32(0))), 1.0])
def test_spiky_closed_form(oracle, q):
    r = np.float32(q * H)
    L = oracle.lib()
    assert L.orc_spiky_w_f(r, H) == pytest.approx(w64(float(r), float(np.float32(H))), rel=2e-6, abs=1e-3)
    assert L.orc_spiky_grad_f(r, H) == pytest.approx(dw64(float(r), float(np.float32(H))), rel=2e-6, abs=1e-1)


def test_spiky_cutoff(oracle):
    L = oracle.lib()
    r_out = float(np.nextafter(np.float32(H), np.float32(1)))  # q > 1
    assert L.orc_spiky_w_f(r_out, H) == 0.0 and L.orc_spiky_grad_f(r_out, H) == 0.0
    assert L.orc_spiky_w_f(H, H) == 0.0  # q == 1 kept, (1-q) = 0


@pytest.mark.parametrize("r", [0.001, 0.004, 0.0075, 0.011])
def test_spiky_gradient_is_h_times_derivative(oracle, r):
    """SpikyKernel::Gradient (Kernel.h:47-57) returns -45/(pi h^3)(1-q)^2, which
    is h * dW/dr (dW/dr = -45/(pi h^4)(1-q)^2): the reference's gradient lacks a
    1/h. The oracle keeps the reference's scaling."""
    L = oracle.lib()
    e = 1e-8
    fd = (L.orc_spiky_w_d(r + e, H) - L.orc_spiky_w_d(r - e, H)) / (2 * e)
    assert L.orc_spiky_grad_d(r, H) == pytest.approx(H * fd, rel=1e-6)


# ------------------------------------------------------------ SummationDensity
def test_scaling_factor_independent_sum(oracle):
    half = int(H / D + 1)
    total = 0.0
    for i in range(-half, half + 1):
        for j in range(-half, half + 1):
            for k in range(-half, half + 1):
                total += D ** 3 * w64(D * math.sqrt(i * i + j * j + k * k))
    assert oracle.lib().orc_scaling_factor_f(D, H) == pytest.approx(1.0 / total, rel=2e-6)
    assert oracle.lib().orc_scaling_factor_d(D, H) == pytest.approx(1.0 / total, rel=1e-12)


def test_perfect_lattice_interior_is_rest_density(oracle, scene):
    p = oracle.make_params()
    x = scene.jittered_lattice((16, 16, 16), (0.3, 0.3, 0.3), D, 0.0, seed=0)
    off, el = oracle.neighbor_query(p, x)
    rho = oracle.density(p, x, off, el)
    idx = np.arange(len(x))
    ijk = np.stack([idx % 16, (idx // 16) % 16, idx // 256], 1)
    interior = np.all((ijk >= 3) & (ijk <= 12), axis=1)
    np.testing.assert_allclose(rho[interior], 1000.0, rtol=2e-5)
    assert np.all(rho[~interior] < 1000.0 * (1 + 2e-5))


def test_isolated_particle_density(oracle):
    p = oracle.make_params()
    x = np.array([[0.2, 0.2, 0.2], [0.6, 0.6, 0.6]], np.float32)
    off, el = oracle.neighbor_query(p, x)
    rho = oracle.density(p, x, off, el)
    m = oracle.lib().orc_scaling_factor_f(D, H) * 1000.0 * D ** 3
    np.testing.assert_allclose(rho, m * 15.0 / (math.pi * H ** 3), rtol=2e-6)


# --------------------------------------------------------------------- lambda
def test_lambda_zero_when_not_compressed(oracle, scene):
    p = oracle.make_params()
    x = scene.jittered_lattice((6, 6, 6), (0.3, 0.3, 0.3), 0.0075, 0.1, seed=3)  # sparse: rho < 1000
    off, el = oracle.neighbor_query(p, x)
    rho = oracle.density(p, x, off, el)
    assert rho.max() < 1000
    lam = oracle.lambdas(p, rho, x, off, el)
    assert np.all(lam == 0.0)


def test_lambda_two_particles(oracle):
    p = oracle.make_params()
    r = 0.006
    x = np.array([[0.5, 0.5, 0.5], [0.5 + r, 0.5, 0.5]], np.float32)
    off, el = oracle.neighbor_query(p, x)
    rho = np.array([1200.0, 1100.0], np.float32)
    lam = oracle.lambdas(p, rho, x, off, el)
    rr = float(x[1, 0] - x[0, 0])
    g2 = dw64(rr) ** 2
    # sum_j |g_j|^2 + |sum_j g_j|^2 with one non-self neighbor = 2 g^2
    np.testing.assert_allclose(lam, [-200.0 / (2 * g2 + 0.1), -100.0 / (2 * g2 + 0.1)], rtol=1e-5)


# --------------------------------------------------------------- displacement
def test_displacement_two_particles_factor_two(oracle):
    p = oracle.make_params()
    r = 0.007
    x = np.array([[0.5, 0.5, 0.5], [0.5, 0.5 + r, 0.5]], np.float32)
    off, el = oracle.neighbor_query(p, x)
    lam = np.array([-3e-9, -1e-9], np.float32)
    dpos = oracle.displacement(p, lam, x, off, el, mode=1)
    rr = float(x[1, 1] - x[0, 1])
    # each endpoint adds dp_ij once and the other endpoint subtracts dp_ji = -dp_ij
    mag = 2 * 10.0 * float(lam.sum()) * dw64(rr)  # times (x_i - x_j)/r
    np.testing.assert_allclose(dpos[0], [0, -mag, 0], rtol=1e-5, atol=1e-15)
    np.testing.assert_allclose(dpos[1], [0, mag, 0], rtol=1e-5, atol=1e-15)
    dpos_atomic = oracle.displacement(p, lam, x, off, el, mode=2)
    np.testing.assert_allclose(dpos_atomic, dpos, rtol=1e-6)


def test_displacement_momentum_and_atomic_emulation(oracle, scene):
    p = oracle.make_params()
    x = scene.jittered_lattice((12, 12, 12), (0.3, 0.3, 0.3), 0.0045, 0.1, seed=4)  # compressed
    off, el = oracle.neighbor_query(p, x)
    rho = oracle.density(p, x, off, el)
    lam = oracle.lambdas(p, rho, x, off, el)
    assert lam.min() < 0
    g = oracle.displacement(p, lam, x, off, el, mode=1).astype(np.float64)
    a = oracle.displacement(p, lam, x, off, el, mode=2).astype(np.float64)
    scale = np.abs(g).max()
    assert np.abs(g.sum(0)).max() < 1e-4 * np.abs(g).sum(0).max()  # pairwise antisymmetry
    np.testing.assert_allclose(g, a, rtol=0, atol=1e-5 * scale)


# ------------------------------------------------------- integrator and step
def test_free_fall_bit_exact(oracle):
    """No pair within h: density stays below 1000, lambda = 0, so the step is
    symplectic Euler alone: v += (0 + g)*dt; x += v*dt in float32."""
    p = oracle.make_params()
    x0 = np.array([[0.2, 0.5, 0.2], [0.5, 0.6, 0.5], [0.8, 0.7, 0.3]], np.float32)
    v0 = np.array([[0.1, 0.0, -0.2], [0.0, 0.3, 0.0], [-0.05, 0.0, 0.07]], np.float32)
    x, v, rho, pairs = oracle.step(p, x0, v0, 20)
    ex, ev = x0.copy(), v0.copy()
    g = np.array([0.0, -9.8, 0.0], np.float32)
    dt = np.float32(0.001)
    for _ in range(20):
        ev = ev + (np.float32(0) + g) * dt
        ex = ex + ev * dt
    np.testing.assert_array_equal(v, ev)
    np.testing.assert_array_equal(x, ex)
    assert pairs == 3
    # and the closed form of the recursion
    t = np.arange(1, 21)
    np.testing.assert_allclose(x[:, 1], x0[:, 1] + v0[:, 1] * 0.02 - 9.8 * 1e-6 * t.sum(), atol=2e-6)


def test_integrator_with_force(oracle):
    p = oracle.make_params()
    x0 = np.array([[0.1, 0.2, 0.3]], np.float32)
    v0 = np.array([[1.0, 2.0, 3.0]], np.float32)
    f = np.array([[5.0, 9.8, -1.0]], np.float32)
    x, v = oracle.integrate(p, x0, v0, f)
    np.testing.assert_allclose(v, [[1.005, 2.0, 2.999]], rtol=1e-6)
    np.testing.assert_allclose(x, x0 + v * 0.001, rtol=1e-7)


def test_pbd_leaves_uncompressed_state(oracle, scene):
    p = oracle.make_params()
    x = scene.jittered_lattice((6, 6, 6), (0.3, 0.3, 0.3), 0.0075, 0.1, seed=3)
    v = np.full_like(x, 0.25)
    off, el = oracle.neighbor_query(p, x)
    x2, v2, rho = oracle.pbd_constrain(p, x, v, off, el)
    np.testing.assert_array_equal(x2, x)
    np.testing.assert_array_equal(v2, v)


def test_step_order_neighbors_before_integration(oracle, scene):
    """The neighbor lists come from the positions before integration
    (PositionBasedFluidModel.cpp:93-94): a fast pair that comes within h
    during the step is not coupled in that step."""
    p = oracle.make_params(g=(0.0, 0.0, 0.0))
    x = np.array([[0.5, 0.5, 0.5], [0.5 + 0.0126, 0.5, 0.5]], np.float32)
    v = np.array([[1.0, 0, 0], [-1.0, 0, 0]], np.float32)
    x1, v1, rho, pairs = oracle.step(p, x, v, 1)
    assert pairs == 2  # self only
    np.testing.assert_allclose(v1, v)


def test_step_is_deterministic_and_thread_independent(oracle, scene):
    p = oracle.make_params()
    x = scene.jittered_lattice((12, 12, 12), (0.3, 0.5, 0.3), D, 0.1, seed=7)
    v = np.zeros_like(x)
    t0 = oracle.set_threads(1)
    a = oracle.step(p, x, v, 2)
    oracle.set_threads(4)
    b = oracle.step(p, x, v, 2)
    oracle.set_threads(max(t0, 1))
    for u, w in zip(a[:3], b[:3]):
        np.testing.assert_array_equal(u, w)


def test_double_step_tracks_float(oracle, scene):
    p32 = oracle.make_params()
    p64 = oracle.make_params(double=True)
    x = scene.jittered_lattice((10, 10, 10), (0.3, 0.5, 0.3), D, 0.1, seed=8)
    v = np.zeros_like(x)
    a = oracle.step(p32, x, v, 3)
    b = oracle.step(p64, x.astype(np.float64), v.astype(np.float64), 3)
    np.testing.assert_allclose(a[0], b[0], atol=1e-6)
    np.testing.assert_allclose(a[2], b[2], rtol=1e-4)

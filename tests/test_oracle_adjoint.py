"""Pins of the oracle's adjoint (step_adjoint / rollout_adjoint), CPU only.

The adjoint holds the neighbor sets fixed; the step is C1 across a set change
(W and its gradient vanish with their first derivative at r = h), so FP64
central differences of the forward step must agree with it. A compressed
lattice (spacing 0.0036 < the 0.005 rest spacing) makes the density exceed
the rest density, so the lambda branch (not only the clamp) is exercised."""
import numpy as np
import pytest


def lattice(scene, dims, spacing, seed):
    return scene.jittered_lattice(dims, (0.3, 0.5, 0.3), spacing, 0.1, seed=seed).astype(np.float64)


def linear_loss(oracle, P, nsteps, cx, cv):
    def L(x, v):
        xs, vs, _, _ = oracle.step(P, x, v, nsteps)
        return float((cx * xs).sum() + (cv * vs).sum())
    return L


def test_adjoint_matches_central_differences(oracle, scene):
    x0 = lattice(scene, (7, 7, 7), 0.0036, 11)
    n = len(x0)
    v0 = 0.05 * (scene.uniform01(12, 3 * n).reshape(n, 3) - 0.5)
    cx = scene.uniform01(13, 3 * n).reshape(n, 3) - 0.5
    cv = 1e-3 * (scene.uniform01(14, 3 * n).reshape(n, 3) - 0.5)
    P = oracle.make_params(double=True)
    _, _, rho, _ = oracle.step(P, x0, v0, 1)
    assert (rho > 1000).mean() > 0.2  # lambda active for a good share of particles
    nsteps = 2
    xb, vb = oracle.rollout_adjoint(P, x0, v0, nsteps, cx, cv)
    L = linear_loss(oracle, P, nsteps, cx, cv)
    rng = np.random.default_rng(0)
    for _ in range(8):
        i, c = int(rng.integers(n)), int(rng.integers(3))
        for base, grad, eps, is_x in ((x0, xb, 1e-7, True), (v0, vb, 1e-5, False)):
            a, b = base.copy(), base.copy()
            a[i, c] += eps
            b[i, c] -= eps
            fd = (L(a, v0) - L(b, v0)) / (2 * eps) if is_x else (L(x0, a) - L(x0, b)) / (2 * eps)
            assert abs(fd - grad[i, c]) <= 1e-4 * abs(fd) + 1e-9, (i, c, is_x, fd, grad[i, c])


def test_adjoint_free_fall_closed_form(oracle, scene):
    """Isolated particles (no neighbor but themselves, density below rest, so
    lambda = 0): x_N = x_0 + N dt v_0 + const, v_N = v_0 + const, hence
    xbar_0 = cx and vbar_0 = N dt cx + cv."""
    x0 = lattice(scene, (4, 4, 4), 0.05, 3)
    n = len(x0)
    v0 = np.zeros_like(x0)
    cx = scene.uniform01(5, 3 * n).reshape(n, 3) - 0.5
    cv = scene.uniform01(6, 3 * n).reshape(n, 3) - 0.5
    P = oracle.make_params(double=True)
    N = 7
    xb, vb = oracle.rollout_adjoint(P, x0, v0, N, cx, cv)
    np.testing.assert_allclose(xb, cx, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(vb, N * 1e-3 * cx + cv, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("nsteps", [1, 3])
def test_adjoint_float_and_double_agree(oracle, scene, nsteps):
    """The FP32 instantiation follows the FP64 one (same algorithm)."""
    x0 = lattice(scene, (6, 6, 6), 0.004, 21)
    n = len(x0)
    v0 = np.zeros_like(x0)
    cx = scene.uniform01(22, 3 * n).reshape(n, 3) - 0.5
    cv = np.zeros_like(cx)
    xd, vd = oracle.rollout_adjoint(oracle.make_params(double=True), x0, v0, nsteps, cx, cv)
    xf, vf = oracle.rollout_adjoint(oracle.make_params(), x0.astype(np.float32), v0.astype(np.float32), nsteps,
                                    cx.astype(np.float32), cv.astype(np.float32))
    assert np.linalg.norm(xf - xd) <= 1e-3 * np.linalg.norm(xd)
    assert np.linalg.norm(vf - vd) <= 1e-3 * np.linalg.norm(vd)

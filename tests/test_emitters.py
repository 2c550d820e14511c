"""Particle emitters and ParticleFluid's copy-in/out (CPU only).

The rotation is pinned by closed forms: the reference's chain
Identity() * q(a, X) * q(b, Y) * q(c, Z) with get3x3Matrix
(quaternion.cpp:24-62,155-171,365-376) reads the quaternion's fields with
different scalar slots, so a rotation input about one axis comes out as a
textbook rotation about another: (a, 0, 0) -> Rz(a), (0, b, 0) -> Ry(-b),
(0, 0, c) -> Rx(-c); every input gives an orthogonal matrix of determinant 1.
The fluid's copy-out -> advance2 -> copy-in cycle (ParticleFluid.cpp:33-124)
runs against a stand-in model whose step is a known map."""
import math
from importlib import import_module

import numpy as np
import pytest
import torch


@pytest.fixture(scope="module")
def em(pkg):
    import __graft_entry__ as G
    return import_module(G.PKG_NAME + ".emitters")


def _rx(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def _ry(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def _rz(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def test_rotation_zero_is_identity(em):
    np.testing.assert_array_equal(em.reference_rotation((0, 0, 0)), np.eye(3, dtype=np.float32))


def test_rotation_half_turn_about_x(em):
    np.testing.assert_allclose(em.reference_rotation((180, 0, 0)), np.diag([-1.0, -1.0, 1.0]), atol=2e-7)


@pytest.mark.parametrize("deg", [30.0, 90.0, -45.0, 200.0])
def test_rotation_single_axis_closed_forms(em, deg):
    a = math.radians(deg)
    np.testing.assert_allclose(em.reference_rotation((deg, 0, 0)), _rz(a), atol=1e-6)
    np.testing.assert_allclose(em.reference_rotation((0, deg, 0)), _ry(-a), atol=1e-6)
    np.testing.assert_allclose(em.reference_rotation((0, 0, deg)), _rx(-a), atol=1e-6)


def test_rotation_is_proper_orthogonal(em):
    rng = np.random.default_rng(5)
    for rot in rng.uniform(-360, 360, size=(20, 3)):
        r = em.reference_rotation(rot).astype(np.float64)
        np.testing.assert_allclose(r @ r.T, np.eye(3), atol=2e-6)
        assert abs(np.linalg.det(r) - 1.0) < 2e-6


def _float_grid(w, d):
    out, x = [], np.float32(-w)
    while x <= np.float32(w):
        out.append(x)
        x = np.float32(x + np.float32(d))
    return out


def test_float_accumulated_grid(em):
    g = em._grid(np.float32(-0.05), np.float32(0.05), np.float32(0.005))
    assert g.dtype == np.float32
    assert g[0] == np.float32(-0.05) and len(g) == len(_float_grid(0.05, 0.005))
    steps = np.diff(g.astype(np.float64)) / 0.005
    np.testing.assert_allclose(steps, 1.0, atol=1e-4)
    assert 20 <= len(g) <= 21


def test_square_emitter_geometry_and_velocity(em):
    e = em.ParticleEmitterSquare(device="cpu", seed=11)
    e.varLocation((0.5, 0.8, 0.25))
    for _ in range(6):
        e.generateParticles()
    p, v = e.gen_pos, e.gen_vel
    assert len(p) > 0 and p.dtype == np.float32
    np.testing.assert_array_equal(p[:, 1], np.float32(0.8))
    local = p.astype(np.float64) - np.array([0.5, 0.8, 0.25])
    assert np.all(np.abs(local[:, [0, 2]]) <= 0.05 + 1e-6)
    k = local[:, [0, 2]] / 0.005
    np.testing.assert_allclose(k, np.round(k), atol=1e-3)
    np.testing.assert_array_equal(v, np.tile(np.float32([0, -1, 0]), (len(p), 1)))


def test_square_emitter_rotated(em):
    e = em.ParticleEmitterSquare(device="cpu", seed=2)
    e.varRotation((90, 0, 0))  # reference input about x: Rz(90) maps (x, 0, z) to (0, x, z)
    e.velocity_magnitude = np.float32(2.0)
    for _ in range(4):
        e.generateParticles()
    assert len(e.gen_pos) > 0
    assert np.all(np.abs(e.gen_pos[:, 0]) < 1e-6)
    np.testing.assert_allclose(e.gen_vel, np.tile([2.0, 0.0, 0.0], (len(e.gen_pos), 1)), atol=1e-6)


def test_round_emitter_inside_disc(em):
    e = em.ParticleEmitterRound(device="cpu", seed=4)
    e.varLocation((0.1, 0.2, 0.3))
    seen = 0
    for _ in range(40):
        e.generateParticles()
        local = e.gen_pos.astype(np.float64) - np.array([0.1, 0.2, 0.3])
        assert np.all(np.hypot(local[:, 0], local[:, 2]) < 0.05 + 1e-6)
        seen += len(e.gen_pos)
    assert seen > 0


def test_draw_rate_and_determinism(em):
    a = em.ParticleEmitterSquare(device="cpu", seed=7)
    b = em.ParticleEmitterSquare(device="cpu", seed=7)
    c = em.ParticleEmitterSquare(device="cpu", seed=8)
    total, frames = 0, 300
    differ = False
    for _ in range(frames):
        for e in (a, b, c):
            e.generateParticles()
        np.testing.assert_array_equal(a.gen_pos, b.gen_pos)
        total += len(a.gen_pos)
        differ |= (a.gen_pos.shape != c.gen_pos.shape) or not np.array_equal(a.gen_pos, c.gen_pos)
    assert differ
    npts = len(_float_grid(0.05, 0.005)) ** 2
    assert a._draws == frames * npts
    # one point in 40 on average; frames with no draw repeat the last emission,
    # so the count is at least the drawn count
    expect = frames * npts / 40.0
    assert total >= 0.85 * expect and total <= 1.25 * expect


def test_empty_draw_repeats_previous_emission(em):
    e = em.ParticleEmitterSquare(device="cpu", seed=1)
    e.width = np.float32(0.0)
    e.height = np.float32(0.0)
    first = None
    for i in range(400):
        e.generateParticles()
        if first is None and len(e.gen_pos):
            first = i
        if first is not None:
            assert len(e.gen_pos) == 1  # one grid point: kept, or the previous emission again
        else:
            assert len(e.gen_pos) == 0
    assert first is not None and e._draws == 400


def test_advance2_appends_after_current(em):
    e = em.ParticleEmitterSquare(device="cpu", seed=3)
    e.varLocation((0.5, 0.5, 0.5))
    cur_x = torch.arange(12, dtype=torch.float32).reshape(4, 3)
    cur_v = -cur_x
    cur_f = cur_x * 10
    e.currentPosition.setValue(cur_x)
    e.currentVelocity.setValue(cur_v)
    e.currentForce.setValue(cur_f)
    while True:
        e.advance2(0.001)
        m = len(e.gen_pos)
        if m:
            break
        assert e.currentPosition.getElementCount() == 4
    x, v, f = (t.getValue() for t in (e.currentPosition, e.currentVelocity, e.currentForce))
    assert x.shape == (4 + m, 3)
    assert torch.equal(x[:4], cur_x) and torch.equal(v[:4], cur_v) and torch.equal(f[:4], cur_f)
    assert torch.equal(x[4:], torch.from_numpy(e.gen_pos))
    assert torch.equal(v[4:], torch.from_numpy(e.gen_vel))
    assert torch.equal(f[4:], torch.from_numpy(e.gen_pos))  # generated forces are the positions


class _ShiftModel:
    """Stand-in for PositionBasedFluidModel: step adds dt to every coordinate
    and 1 to every velocity component."""

    def __init__(self, fluid_mod):
        self.m_position = fluid_mod.ArrayField()
        self.m_velocity = fluid_mod.ArrayField()
        self.m_forceDensity = fluid_mod.ArrayField()
        self.steps = []
        self.ctx = self
        self.device = torch.device("cpu")

    def params(self):
        return {"dt": 0.001}

    def step(self, dt=None):
        self.steps.append((dt, self.m_forceDensity.getValue().abs().sum().item()))
        self.m_position.setValue(self.m_position.getValue() + dt)
        self.m_velocity.setValue(self.m_velocity.getValue() + 1)


def test_fluid_copy_out_in_cycle(em, fluid):
    model = _ShiftModel(fluid)
    pf = fluid.ParticleFluid(model=model)
    e1 = em.ParticleEmitterSquare(device="cpu", seed=21)
    e2 = em.ParticleEmitterRound(device="cpu", seed=22)
    e1.varLocation((0.2, 0.7, 0.2))
    e2.varLocation((0.7, 0.7, 0.7))
    pf.addParticleEmitter(e1)
    pf.addParticleEmitter(e2)
    pf.resetStatus()
    assert pf.currentPosition.getElementCount() == 0
    n1 = n2 = 0
    for frame in range(30):
        x_before = pf.currentPosition.getValue().clone()
        v_before = pf.currentVelocity.getValue().clone()
        stepped = pf.advance(0.002)
        g1, g2 = len(e1.gen_pos), len(e2.gen_pos)
        # emitter segments: e1's old, e1's new, e2's old, e2's new
        parts_x = [x_before[:n1], torch.from_numpy(e1.gen_pos), x_before[n1:n1 + n2], torch.from_numpy(e2.gen_pos)]
        parts_v = [v_before[:n1], torch.from_numpy(e1.gen_vel), v_before[n1:n1 + n2], torch.from_numpy(e2.gen_vel)]
        n1, n2 = n1 + g1, n2 + g2
        assert stepped == n1 + n2 == pf.currentPosition.getElementCount()
        if stepped:
            assert torch.equal(pf.currentPosition.getValue(), torch.cat(parts_x) + 0.002)
            assert torch.equal(pf.currentVelocity.getValue(), torch.cat(parts_v) + 1)
            assert model.steps[-1] == (0.002, 0.0)  # forces zeroed before the step
        assert e1.currentPosition.getElementCount() == n1 and e2.currentPosition.getElementCount() == n2
    assert n1 > 0 and n2 > 0 and len(model.steps) > 0


def test_fluid_without_emitters_steps_its_particles(fluid):
    model = _ShiftModel(fluid)
    pf = fluid.ParticleFluid(model=model)
    assert pf.advance(0.001) == 0 and not model.steps
    model.m_position.setValue(torch.zeros(5, 3))
    model.m_velocity.setValue(torch.zeros(5, 3))
    model.m_forceDensity.setValue(torch.ones(5, 3))
    assert pf.advance(0.001) == 5
    assert torch.equal(pf.currentPosition.getValue(), torch.full((5, 3), 0.001))
    pf.self_update = False
    assert pf.advance(0.001) == 0 and len(model.steps) == 1

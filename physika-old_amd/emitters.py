"""Particle emitters: the sources that feed a ParticleFluid a variable number
of particles (Dynamics/ParticleSystem/ParticleEmitter.cpp,
ParticleEmitterSquare.cpp, ParticleEmitterRound.cpp).

An emitter owns the state of the particles it has emitted between frames:
ParticleFluid.advance copies each emitter's segment of the fluid state back
into the emitter, lets every emitter append its new particles (advance2), and
concatenates the emitters' arrays into the fluid again before the step
(ParticleFluid.cpp:33-116).

Kept as in the reference:
- the sampling grid, walked by float32 accumulation (x += d while x <= w);
- the emitted velocity magnitude * R * (0, -1, 0);
- R from the reference's quaternion chain Identity() * q(rx, X) * q(ry, Y) *
  q(rz, Z) and get3x3Matrix (quaternion.cpp:24-62,155-171,365-376). That code
  multiplies with the first field as the scalar part but builds the matrix
  with the last one as the scalar, and Identity() is (0, 0, 0, 1): for zero
  rotation R is the identity; a rotation of 180 degrees about x gives
  diag(-1, -1, 1). R is computed with these same formulas, in float32;
- a frame in which no point is drawn emits the previous frame's particles
  again (generateParticles keeps gen_pos when the draw is empty);
- the forces of emitted particles are their positions (advance2 copies gen_pos
  into the force array); the fluid's integrator zeroes forces before use.

Deviation: the reference keeps a grid point when rand() % 40 == 0, with libc's
rand seeded by the wall clock; here the draw is SplitMix64 of the emitter's
seed and a running counter (output % 40 == 0), so runs are reproducible.
"""
from __future__ import annotations

import math
import numpy as np
import torch

from .fluid import ArrayField
from .scene import splitmix64


def _quat_axis_angle(angle: np.float32, axis) -> np.ndarray:
    """Quaternion(angle, axis) (quaternion.cpp:52-62): fields (x, y, z, w) =
    (cos(a/2), axis * sin(a/2) / |axis|), float32."""
    f = np.float32
    c = f(math.cos(f(0.5) * angle))
    s = f(math.sin(f(0.5) * angle))
    ax = np.asarray(axis, np.float32)
    n = f(math.sqrt(f(f(ax[0] * ax[0]) + f(ax[1] * ax[1])) + f(ax[2] * ax[2])))
    t = f(s / n)
    return np.array([c, f(ax[0] * t), f(ax[1] * t), f(ax[2] * t)], np.float32)


def _quat_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Quaternion::operator* (quaternion.cpp:155-171), float32."""
    f = np.float32
    x1, y1, z1, w1 = (f(v) for v in a)
    x2, y2, z2, w2 = (f(v) for v in b)
    return np.array([f(f(f(x1 * x2) - f(y1 * y2)) - f(z1 * z2)) - f(w1 * w2),
                     f(f(f(x1 * y2) + f(y1 * x2)) + f(z1 * w2)) - f(w1 * z2),
                     f(f(f(x1 * z2) + f(z1 * x2)) + f(w1 * y2)) - f(y1 * w2),
                     f(f(f(x1 * w2) + f(w1 * x2)) + f(y1 * z2)) - f(z1 * y2)], np.float32)


def _quat_matrix(q: np.ndarray) -> np.ndarray:
    """Quaternion::get3x3Matrix (quaternion.cpp:365-376), float32."""
    f = np.float32
    x, y, z, w = (f(v) for v in q)
    x2, y2, z2 = f(x + x), f(y + y), f(z + z)
    xx, yy, zz = f(x2 * x), f(y2 * y), f(z2 * z)
    xy, xz, xw = f(x2 * y), f(x2 * z), f(x2 * w)
    yz, yw, zw = f(y2 * z), f(y2 * w), f(z2 * w)
    one = f(1)
    return np.array([[f(f(one - yy) - zz), f(xy - zw), f(xz + yw)],
                     [f(xy + zw), f(f(one - xx) - zz), f(yz - xw)],
                     [f(xz - yw), f(yz + xw), f(f(one - xx) - yy)]], np.float32)


def reference_rotation(rotation_deg) -> np.ndarray:
    """The emitter's rotation matrix for Node::varRotation (degrees), as
    ParticleEmitterSquare.cpp:37-45 forms it."""
    f = np.float32
    rad = [f(float(f(f(r) / f(180.0))) * math.pi) for r in rotation_deg]  # float / 180.0f * M_PI
    q = np.array([0, 0, 0, 1], np.float32)  # Quaternion::Identity()
    for angle, axis in zip(rad, ((1, 0, 0), (0, 1, 0), (0, 0, 1))):
        q = _quat_mul(q, _quat_axis_angle(angle, axis))
    return _quat_matrix(q)


def _matvec(m: np.ndarray, p: np.ndarray) -> np.ndarray:
    """SquareMatrix * Vector in float32, rows as ((m0 p0 + m1 p1) + m2 p2)."""
    f = np.float32
    return np.array([f(f(f(m[i, 0] * p[0]) + f(m[i, 1] * p[1])) + f(m[i, 2] * p[2])) for i in range(3)], np.float32)


class ParticleEmitter:
    """ParticleEmitter (ParticleEmitter.h:19-45): location and rotation (Node
    variables), velocity magnitude 1, sampling distance 0.005; the emitted
    particles' state in currentPosition / currentVelocity / currentForce."""

    def __init__(self, name: str = "particle emitter", device="cuda:0", seed: int = 0):
        self.name = name
        self.device = torch.device(device)
        self.location = np.zeros(3, np.float32)
        self.rotation = np.zeros(3, np.float32)
        self.velocity_magnitude = np.float32(1.0)
        self.sampling_distance = np.float32(0.005)
        self.seed = int(seed)
        self._draws = 0  # position in the emitter's SplitMix64 stream
        self.gen_pos = np.zeros((0, 3), np.float32)
        self.gen_vel = np.zeros((0, 3), np.float32)
        self.currentPosition = ArrayField()
        self.currentVelocity = ArrayField()
        self.currentForce = ArrayField()

    # Node variables
    def varLocation(self, value=None):
        if value is not None:
            self.location = np.asarray(value, np.float32)
        return self.location

    def varRotation(self, value=None):
        if value is not None:
            self.rotation = np.asarray(value, np.float32)
        return self.rotation

    def _sampling(self) -> np.float32:
        d = np.float32(self.sampling_distance)
        return d if d >= np.float32(1e-6) else np.float32(0.005)  # EPSILON guard of generateParticles

    def _keep(self, count: int) -> np.ndarray:
        r = splitmix64(self.seed, count, self._draws)
        self._draws += count
        return (r % np.uint64(40)) == 0

    def _emit(self, points, inside) -> None:
        """Keeps the drawn points of the local grid, rotated and placed at the
        location; an empty draw leaves the previous frame's particles."""
        pts = np.asarray(points, np.float32).reshape(-1, 3)
        keep = np.zeros(len(pts), bool)
        keep[inside] = self._keep(int(inside.sum()))  # one draw per candidate point, as rand() is reached
        if not keep.any():
            return
        rot = reference_rotation(self.rotation)
        v0 = _matvec(rot * np.float32(self.velocity_magnitude), np.array([0, -1, 0], np.float32))
        q = np.stack([_matvec(rot, p) for p in pts[keep]]).astype(np.float32)
        self.gen_pos = (q + self.location).astype(np.float32)
        self.gen_vel = np.tile(v0, (len(q), 1)).astype(np.float32)

    def generateParticles(self) -> None:
        pass

    def advance2(self, dt: float) -> None:
        """ParticleEmitter::advance2 (ParticleEmitter.cpp:28-72): emitted
        particles appended after the current ones."""
        self.generateParticles()
        cur_x = self.currentPosition.getValue()
        n = 0 if cur_x is None else cur_x.shape[0]
        m = len(self.gen_pos)
        if n + m == 0:
            return
        gp = torch.from_numpy(self.gen_pos).to(self.device)
        gv = torch.from_numpy(self.gen_vel).to(self.device)
        if n:
            self.currentPosition.setValue(torch.cat([cur_x, gp]))
            self.currentVelocity.setValue(torch.cat([self.currentVelocity.getValue(), gv]))
            self.currentForce.setValue(torch.cat([self.currentForce.getValue(), gp.clone()]))
        else:
            self.currentPosition.setValue(gp)
            self.currentVelocity.setValue(gv)
            self.currentForce.setValue(gp.clone())


def _grid(lo: np.float32, hi: np.float32, d: np.float32) -> np.ndarray:
    """for (Real x = lo; x <= hi; x += d) in float32."""
    out = []
    x = np.float32(lo)
    while x <= hi:
        out.append(x)
        x = np.float32(x + d)
    return np.array(out, np.float32)


class ParticleEmitterSquare(ParticleEmitter):
    """ParticleEmitterSquare (ParticleEmitterSquare.cpp:25-68): grid points
    (x, 0, z), |x| <= width, |z| <= height (defaults 0.05)."""

    def __init__(self, name: str = "square emitter", device="cuda:0", seed: int = 0):
        super().__init__(name, device, seed)
        self.width = np.float32(0.05)
        self.height = np.float32(0.05)

    def generateParticles(self) -> None:
        d = self._sampling()
        xs = _grid(-self.width, self.width, d)
        zs = _grid(-self.height, self.height, d)
        pts = [(x, np.float32(0), z) for x in xs for z in zs]
        self._emit(pts, np.ones(len(pts), bool))


class ParticleEmitterRound(ParticleEmitter):
    """ParticleEmitterRound (ParticleEmitterRound.cpp:25-70): grid points
    (x, 0, y) of the square [-r, r]^2 with |p| < r (default r = 0.05)."""

    def __init__(self, name: str = "round emitter", device="cuda:0", seed: int = 0):
        super().__init__(name, device, seed)
        self.radius = np.float32(0.05)

    def generateParticles(self) -> None:
        d = self._sampling()
        r = np.float32(self.radius)
        g = _grid(-r, r, d)
        pts = [(x, np.float32(0), y) for x in g for y in g]
        inside = np.array([np.sqrt(np.float32(np.float32(x * x) + np.float32(0)) + np.float32(y * y)) < r
                           for x, _, y in pts], bool)
        self._emit(pts, inside)

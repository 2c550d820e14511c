"""MI355X-native PBD particle-fluid step of PhysIKA (NeighborQuery, DensityPBD,
ParticleIntegrator) behind the reference's module interface.

The compute path is ``libpbf.so`` (csrc/): hand-written HIP kernels for gfx950
behind the C ABI declared in include/pbf.h. This package is the Python face of
that library: ``_capi`` binds the C ABI with ctypes, ``fluid`` mirrors the
reference's module classes, ``scene`` generates seeded inputs.

The directory name contains a hyphen, so it is imported by path (see
``__graft_entry__.load_package``) under the module name ``physika_old_amd``.
"""
from . import scene  # noqa: F401

__all__ = ["scene"]

"""MI355X-native diomedes physics step (HIP kernels for gfx950 behind a C ABI).

The compute lives in ``libdiomedes_hip.so`` (``diomedes_amd/csrc``), built in
tree by ``python -m diomedes_amd.build``; this package is its host-side
binding, mirroring the reference engine's physics interface
(``physics.AvbdSolver``, ``physics.PhysicsContext``) and the seeded workloads
(``scenes``).
"""

from . import records
from .physics import (
    AvbdBody,
    AvbdContainer,
    AvbdOptions,
    AvbdSolver,
    BallSocket,
    Bend,
    HardBallSocket,
    HardBend,
    HardRod,
    HardRope,
    PhysicsContext,
    Rod,
    Rope,
    Scene,
    Shape,
    Spring,
    constraints_from_avbd,
    decode_contacts,
    gpu_bodies,
    state_from_bodies,
)

__all__ = [
    "records",
    "AvbdBody",
    "AvbdContainer",
    "AvbdOptions",
    "AvbdSolver",
    "BallSocket",
    "Bend",
    "HardBallSocket",
    "HardBend",
    "HardRod",
    "HardRope",
    "PhysicsContext",
    "Rod",
    "Rope",
    "Scene",
    "Shape",
    "Spring",
    "constraints_from_avbd",
    "decode_contacts",
    "gpu_bodies",
    "state_from_bodies",
]

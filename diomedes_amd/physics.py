"""Host-side mirror of the reference engine's physics interface.

Names, fields and defaults follow engine/src/physics/{gpu,shape,constraints}.rs
and the six Renderer physics methods (engine/src/render/renderer.rs:75-175);
the work runs in the HIP library through its C ABI:

- ``AvbdBody``, ``AvbdContainer``, ``AvbdOptions``, ``Spring`` / ``Rod`` /
  ``BallSocket`` / ``Rope`` and ``state_from_bodies`` / ``constraints_from_avbd``
  build the reference's records;
- ``PhysicsContext`` exposes the Renderer methods (``broad_phase_submit``,
  ``broad_phase_count``, ``narrow_phase_submit``, ``narrow_phase_read``,
  ``gpu_physics_submit``, ``gpu_physics_read``) and the fused device-resident
  world (``upload``, ``step``, ``download``);
- ``AvbdSolver`` keeps the reference's world object (bodies, container,
  constraints) and steps it on the GPU.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Iterable, Optional, Sequence

import numpy as np

from . import _lib
from . import records as R

DEFAULT_ROD_STIFFNESS = 1000.0  # physics/constraints.rs:5-7
DEFAULT_SOCKET_STIFFNESS = 1000.0
DEFAULT_ROPE_STIFFNESS = 1000.0


# ---------------------------------------------------------------- shapes
@dataclass(frozen=True)
class Shape:
    """physics/shape.rs:7-15 plus the BOX / HULL extensions."""

    tag: int
    params: tuple

    @staticmethod
    def sphere(radius: float) -> "Shape":
        return Shape(R.SPHERE, (float(radius),))

    @staticmethod
    def cube(half_extent: float) -> "Shape":
        return Shape(R.CUBE, (float(half_extent),))

    @staticmethod
    def tetrahedron(corners: Sequence[Sequence[float]]) -> "Shape":
        return Shape(R.TETRAHEDRON, tuple(tuple(map(float, c)) for c in corners))

    @staticmethod
    def box(half_extents: Sequence[float]) -> "Shape":
        return Shape(R.BOX, tuple(map(float, half_extents)))

    @staticmethod
    def hull(first: int, count: int) -> "Shape":
        return Shape(R.HULL, (int(first), int(count)))

    def packed(self) -> np.ndarray:
        """GpuShape corners (compute.rs:41-72)."""
        c = np.zeros((4, 4), np.float32)
        if self.tag == R.SPHERE:
            c[0, 0] = self.params[0]
        elif self.tag == R.CUBE:
            c[0, :3] = self.params[0]
        elif self.tag == R.TETRAHEDRON:
            for i, p in enumerate(self.params):
                c[i, :3] = p
        elif self.tag == R.BOX:
            c[0, :3] = self.params
        elif self.tag == R.HULL:
            c[0, 0], c[0, 1] = self.params
        return c


def _f(x) -> np.float32:
    return np.float32(x)


@dataclass
class AvbdBody:
    """physics/gpu.rs:19-99 (inertia rules of the constructors kept in f32)."""

    shape: Shape
    position: tuple = (0.0, 0.0, 0.0)
    orientation: tuple = (0.0, 0.0, 0.0, 1.0)  # x, y, z, w
    velocity: tuple = (0.0, 0.0, 0.0)
    angular_velocity: tuple = (0.0, 0.0, 0.0)
    inv_mass: float = 0.0
    inv_moment: tuple = (0.0, 0.0, 0.0)
    friction: float = 0.8
    prev_velocity: tuple = (0.0, 0.0, 0.0)

    @staticmethod
    def sphere(position, radius: float, mass: float) -> "AvbdBody":
        m, r = _f(mass), _f(radius)
        inv_mass = _f(1.0) / m if m > 0 else _f(0.0)
        inv_moment = _f(1.0) / (_f(0.4) * m * r * r) if m > 0 else _f(0.0)
        return AvbdBody(Shape.sphere(r), tuple(position), inv_mass=float(inv_mass), inv_moment=(float(inv_moment),) * 3)

    @staticmethod
    def cube(position, half_extent: float, mass: float) -> "AvbdBody":
        m, h = _f(mass), _f(half_extent)
        inv_mass = _f(1.0) / m if m > 0 else _f(0.0)
        inv_moment = _f(1.0) / (_f(2.0 / 3.0) * m * h * h) if m > 0 else _f(0.0)
        return AvbdBody(Shape.cube(h), tuple(position), inv_mass=float(inv_mass), inv_moment=(float(inv_moment),) * 3)

    @staticmethod
    def tetrahedron(position, corners, mass: float) -> "AvbdBody":
        m = _f(mass)
        radius = _f(0.0)
        for c in corners:
            c = np.asarray(c, np.float32)
            radius = max(radius, np.sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]))
        inv_mass = _f(1.0) / m if m > 0 else _f(0.0)
        inv_moment = _f(1.0) / (_f(0.4) * m * radius * radius) if m > 0 else _f(0.0)
        return AvbdBody(
            Shape.tetrahedron(corners), tuple(position), inv_mass=float(inv_mass), inv_moment=(float(inv_moment),) * 3
        )

    @staticmethod
    def box(position, half_extents, mass: float) -> "AvbdBody":
        """Extension: solid box, I = m/3 (hy^2 + hz^2) etc. for half extents h."""
        m = _f(mass)
        h = np.asarray(half_extents, np.float32)
        inv_mass = _f(1.0) / m if m > 0 else _f(0.0)
        if m > 0:
            inv_moment = tuple(
                float(_f(1.0) / (m / _f(3.0) * (h[j] * h[j] + h[k] * h[k])))
                for j, k in ((1, 2), (0, 2), (0, 1))
            )
        else:
            inv_moment = (0.0, 0.0, 0.0)
        return AvbdBody(Shape.box(h), tuple(position), inv_mass=float(inv_mass), inv_moment=inv_moment)

    def is_static(self) -> bool:
        return self.inv_mass == 0.0


@dataclass
class AvbdContainer:
    """physics/gpu.rs:117-123: a static box whose interior is free space."""

    center: tuple = (0.0, 0.0, 0.0)
    rotation: tuple = (0.0, 0.0, 0.0, 1.0)
    half_extent: float = 1.0

    def to_c(self) -> R.Container:
        c = R.Container()
        c.center[:] = [float(v) for v in self.center]
        c.rotation[:] = [float(v) for v in self.rotation]
        c.half_extent = float(self.half_extent)
        c.enabled = 1
        return c


@dataclass
class AvbdOptions:
    """physics/gpu.rs:127-155 (defaults of AvbdOptions::default)."""

    dt: float = 1.0 / 60.0
    gravity: tuple = (0.0, -10.0, 0.0)
    iterations: int = 10
    alpha: float = 0.99
    beta_lin: float = 10000.0
    margin: float = 0.01
    penalty_min: float = 1.0
    penalty_max: float = 1e10
    gamma: float = 0.999
    friction: float = 0.8

    def to_c(self) -> R.Options:
        o = R.Options()
        o.dt = self.dt
        o.gravity[:] = [float(v) for v in self.gravity]
        o.iterations = int(self.iterations)
        o.alpha, o.beta_lin, o.margin = self.alpha, self.beta_lin, self.margin
        o.penalty_min, o.penalty_max = self.penalty_min, self.penalty_max
        o.gamma, o.friction = self.gamma, self.friction
        return o

    def run_options(self) -> R.RunOptions:
        """AvbdRunOptions as the demos derive it (demos/avbd_box/src/main.rs:51-58)."""
        r = R.RunOptions()
        r.dt = self.dt
        r.gravity[:] = [float(v) for v in self.gravity]
        r.alpha, r.beta_lin, r.penalty_max = self.alpha, self.beta_lin, self.penalty_max
        r.iterations = int(self.iterations)
        return r


# ------------------------------------------------------------ constraints
@dataclass
class Spring:
    body_a: int
    body_b: int
    anchor_a: tuple
    anchor_b: tuple
    stiffness: float
    rest_length: float
    kind = R.SPRING
    name = "spring"


@dataclass
class Rod:
    body_a: int
    body_b: int
    anchor_a: tuple
    anchor_b: tuple
    rest_length: float
    stiffness: float = DEFAULT_ROD_STIFFNESS
    kind = R.ROD
    name = "rod"


@dataclass
class BallSocket:
    body_a: int
    body_b: int
    anchor_a: tuple
    anchor_b: tuple
    stiffness: float = DEFAULT_SOCKET_STIFFNESS
    kind = R.BALL_SOCKET
    name = "ball and socket"


@dataclass
class Rope:
    body_a: int
    body_b: int
    anchor_a: tuple
    anchor_b: tuple
    max_length: float
    stiffness: float = DEFAULT_ROPE_STIFFNESS
    kind = R.ROPE
    name = "rope"


@dataclass
class HardRod:
    """A rod held at rest_length by a multiplier (no counterpart in the
    reference, whose Rod is a stiff spring)."""
    body_a: int
    body_b: int
    anchor_a: tuple
    anchor_b: tuple
    rest_length: float
    kind = R.HARD_ROD
    name = "hard rod"
    stiffness = 0.0


@dataclass
class HardBallSocket:
    """Two anchors held together by three multipliers."""
    body_a: int
    body_b: int
    anchor_a: tuple
    anchor_b: tuple
    kind = R.HARD_SOCKET
    name = "hard ball and socket"
    stiffness = 0.0


@dataclass
class HardRope:
    """At most max_length between the anchors, held by a multiplier that only
    pulls."""
    body_a: int
    body_b: int
    anchor_a: tuple
    anchor_b: tuple
    max_length: float
    kind = R.HARD_ROPE
    name = "hard rope"
    stiffness = 0.0


def rest_rotation(orientation_a, orientation_b) -> tuple:
    """The rest rotation of a bending joint that holds two bodies at their
    current relative orientation: conj(q_a) * q_b in float32 (x, y, z, w)."""
    ax, ay, az, aw = (np.float32(v) for v in orientation_a)
    bx, by, bz, bw = (np.float32(v) for v in orientation_b)
    ax, ay, az = -ax, -ay, -az
    return (float(aw * bx + ax * bw + ay * bz - az * by), float(aw * by - ax * bz + ay * bw + az * bx),
            float(aw * bz + ax * by - ay * bx + az * bw), float(aw * bw - ax * bx - ay * by - az * bz))


@dataclass
class Bend:
    """Holds b's orientation at a's times `rest` with an angular spring of
    the given stiffness (three angular rows; no counterpart in the reference,
    whose ropes do not bend)."""
    body_a: int
    body_b: int
    rest: tuple = (0.0, 0.0, 0.0, 1.0)
    stiffness: float = DEFAULT_ROD_STIFFNESS
    kind = R.BEND
    name = "bend"


@dataclass
class HardBend:
    """Holds b's orientation at a's times `rest` with three multipliers."""
    body_a: int
    body_b: int
    rest: tuple = (0.0, 0.0, 0.0, 1.0)
    kind = R.HARD_BEND
    name = "hard bend"
    stiffness = 0.0


def constraints_from_avbd(constraints: Iterable) -> np.ndarray:
    """gpu_physics.rs:149-193: anchor_a.w = rest / max length (0 for sockets),
    anchor_b.w = stiffness."""
    cons = list(constraints)
    out = np.zeros(len(cons), R.AVBD_CONSTRAINT)
    for i, c in enumerate(cons):
        out[i]["a"], out[i]["b"], out[i]["kind"] = c.body_a, c.body_b, c.kind
        if isinstance(c, (Bend, HardBend)):
            out[i]["anchor_a"][:] = c.rest
            out[i]["anchor_b"][3] = c.stiffness
            continue
        out[i]["anchor_a"][:3] = c.anchor_a
        out[i]["anchor_b"][:3] = c.anchor_b
        if isinstance(c, (Spring, Rod, HardRod)):
            out[i]["anchor_a"][3] = c.rest_length
        elif isinstance(c, (Rope, HardRope)):
            out[i]["anchor_a"][3] = c.max_length
        out[i]["anchor_b"][3] = c.stiffness
    return out


def gpu_bodies(bodies: Sequence[AvbdBody]) -> np.ndarray:
    """GpuBody records (demos/avbd_box/src/main.rs:232-243)."""
    out = np.zeros(len(bodies), R.GPU_BODY)
    for i, b in enumerate(bodies):
        out[i]["position"][:3] = b.position
        out[i]["orientation"] = b.orientation
        out[i]["tag"] = b.shape.tag
        out[i]["corners"] = b.shape.packed()
    return out


def state_from_bodies(bodies: Sequence[AvbdBody]) -> np.ndarray:
    """gpu_physics.rs:106-119."""
    out = np.zeros(len(bodies), R.BODY_STATE)
    for i, b in enumerate(bodies):
        out[i]["vel"][:3] = b.velocity
        out[i]["ang"][:3] = b.angular_velocity
        out[i]["prev_vel"][:3] = b.prev_velocity
        out[i]["inv_moment"][:3] = b.inv_moment
        out[i]["inv_mass"] = b.inv_mass
        out[i]["friction"] = b.friction
    return out


# ---------------------------------------------------------------- scenes
@dataclass
class Scene:
    """A world as arrays of the reference's records, ready for upload."""

    bodies: np.ndarray  # GPU_BODY
    state: np.ndarray  # BODY_STATE
    constraints: np.ndarray = field(default_factory=lambda: np.zeros(0, R.AVBD_CONSTRAINT))
    container: Optional[AvbdContainer] = None
    hull: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))
    options: AvbdOptions = field(default_factory=AvbdOptions)
    name: str = "scene"

    @property
    def n(self) -> int:
        return len(self.bodies)

    @property
    def n_dynamic(self) -> int:
        return int(np.count_nonzero(self.state["inv_mass"] != 0))

    @staticmethod
    def from_bodies(bodies: Sequence[AvbdBody], constraints=(), container=None, options=None, name="scene"):
        return Scene(
            gpu_bodies(bodies),
            state_from_bodies(bodies),
            constraints_from_avbd(constraints),
            container,
            np.zeros((0, 3), np.float32),
            options or AvbdOptions(),
            name,
        )


# ------------------------------------------------------------- decoded narrow
def decode_contacts(raw: np.ndarray) -> np.ndarray:
    """compute.rs:738-750: the valid slots, in slot order."""
    return raw[raw["valid"] != 0]


class PhysicsContext:
    """One GPU physics context (a dio_ctx): the reference's NarrowPhaseCompute
    and AvbdCompute behind the Renderer's physics methods, plus the fused
    device-resident world."""

    def __init__(self, max_bodies: int, max_pairs: int = 0, max_contacts: int = 0, max_constraints: int = 0,
                 max_hull_vertices: int = 0, device: int = 0):
        self._lib = _lib.load()
        lim = R.Limits(max_bodies, max_pairs, max_contacts, max_constraints, max_hull_vertices)
        h = ctypes.c_void_p()
        _lib.check(None, self._lib.dio_create(device, ctypes.byref(lim), ctypes.byref(h)))
        self._h = h
        self.max_bodies = max_bodies
        self._n = 0
        self._m = 0

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.dio_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, status: int) -> None:
        _lib.check(self._h, status)

    def synchronize(self) -> None:
        """Wait for everything enqueued on the context's stream."""
        self._check(self._lib.dio_synchronize(self._h))

    # ---- Renderer::broad_phase_submit / broad_phase_count
    def broad_phase_submit(self, bodies: np.ndarray) -> None:
        bodies = np.ascontiguousarray(bodies, R.GPU_BODY)
        self._check(self._lib.dio_broad_submit(self._h, _lib.ptr(bodies), len(bodies)))

    def broad_phase_count(self) -> int:
        n = ctypes.c_uint32()
        self._check(self._lib.dio_broad_count(self._h, ctypes.byref(n)))
        return n.value

    def broad_phase_pairs(self) -> np.ndarray:
        count = self.broad_phase_count()
        out = np.zeros(count, R.GPU_PAIR)
        n = ctypes.c_uint32()
        self._check(self._lib.dio_broad_read_pairs(self._h, _lib.ptr(out), count, ctypes.byref(n)))
        return out

    def broad_phase_aabbs(self, n: int) -> np.ndarray:
        out = np.zeros((n, 8), np.float32)
        self._check(self._lib.dio_broad_read_aabbs(self._h, _lib.ptr(out, ctypes.c_float), n))
        return out

    # ---- Renderer::narrow_phase_submit / narrow_phase_read
    def narrow_phase_submit(self, pair_count: int) -> None:
        self._check(self._lib.dio_narrow_submit(self._h, pair_count))

    def narrow_phase_run(self, bodies: np.ndarray, pairs: np.ndarray) -> np.ndarray:
        """NarrowPhaseCompute::run (compute.rs:697-705): raw slots for caller pairs."""
        bodies = np.ascontiguousarray(bodies, R.GPU_BODY)
        pairs = np.ascontiguousarray(pairs, R.GPU_PAIR)
        self._check(self._lib.dio_narrow_submit_pairs(self._h, _lib.ptr(bodies), len(bodies), _lib.ptr(pairs),
                                                      len(pairs)))
        return self.narrow_phase_raw()

    def narrow_phase_raw(self) -> np.ndarray:
        n = ctypes.c_uint32()
        # query the slot count, then read
        status = self._lib.dio_narrow_read(self._h, None, 0, ctypes.byref(n))
        if status not in (0, 1):
            self._check(status)
        out = np.zeros(n.value, R.GPU_CONTACT_OUT)
        self._check(self._lib.dio_narrow_read(self._h, _lib.ptr(out), n.value, ctypes.byref(n)))
        return out

    def narrow_phase_read(self) -> np.ndarray:
        return decode_contacts(self.narrow_phase_raw())

    # ---- Renderer::gpu_physics_submit / gpu_physics_read
    def gpu_physics_submit(self, body_state, positions, orientations, contacts, constraints, offsets, indices,
                           options) -> None:
        st = np.ascontiguousarray(body_state, R.BODY_STATE)
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        rot = np.ascontiguousarray(orientations, np.float32).reshape(-1, 4)
        con = np.ascontiguousarray(contacts, R.AVBD_CONTACT)
        cns = np.ascontiguousarray(constraints, R.AVBD_CONSTRAINT)
        off = np.ascontiguousarray(offsets, np.uint32)
        idx = np.ascontiguousarray(indices, np.uint32)
        ro = options.run_options() if isinstance(options, AvbdOptions) else options
        self._n, self._m = len(st), len(con)
        self._check(
            self._lib.dio_avbd_submit(
                self._h, _lib.ptr(st), _lib.ptr(pos, ctypes.c_float), _lib.ptr(rot, ctypes.c_float), len(st),
                _lib.ptr(con), len(con), _lib.ptr(cns), len(cns), _lib.ptr(off, ctypes.c_uint32),
                _lib.ptr(idx, ctypes.c_uint32), len(idx), ctypes.byref(ro),
            )
        )

    def gpu_physics_read(self) -> dict:
        """AvbdGpuResult (gpu_physics.rs:86-94) as arrays."""
        n, m = self._n, self._m
        out = {
            "positions": np.zeros((n, 3), np.float32),
            "orientations": np.zeros((n, 4), np.float32),
            "velocities": np.zeros((n, 3), np.float32),
            "angular_velocities": np.zeros((n, 3), np.float32),
            "prev_velocities": np.zeros((n, 3), np.float32),
            "lambda": np.zeros((m, 3), np.float32),
            "penalty": np.zeros((m, 3), np.float32),
        }
        fp = lambda a: _lib.ptr(a, ctypes.c_float)  # noqa: E731
        self._check(
            self._lib.dio_avbd_read(self._h, fp(out["positions"]), fp(out["orientations"]), fp(out["velocities"]),
                                    fp(out["angular_velocities"]), fp(out["prev_velocities"]), fp(out["lambda"]),
                                    fp(out["penalty"]))
        )
        return out

    # ---- fused world
    def upload(self, scene: Scene) -> None:
        bodies = np.ascontiguousarray(scene.bodies, R.GPU_BODY)
        state = np.ascontiguousarray(scene.state, R.BODY_STATE)
        cons = np.ascontiguousarray(scene.constraints, R.AVBD_CONSTRAINT)
        if len(scene.hull):
            hull = np.ascontiguousarray(scene.hull, np.float32)
            self._check(self._lib.dio_world_upload_hulls(self._h, _lib.ptr(hull, ctypes.c_float), len(hull)))
        ct = scene.container.to_c() if scene.container is not None else None
        self._check(
            self._lib.dio_world_upload(self._h, _lib.ptr(bodies), _lib.ptr(state), len(bodies), _lib.ptr(cons),
                                       len(cons), ctypes.byref(ct) if ct is not None else None)
        )
        self._n = len(bodies)

    def batch_upload(self, batch) -> None:
        """dio_batch_upload: a scene table (``scenes.BatchedScenes`` or any
        object with bodies, state, offsets, containers, hull and optional
        constraints in global body indices)."""
        bodies = np.ascontiguousarray(batch.bodies, R.GPU_BODY)
        state = np.ascontiguousarray(batch.state, R.BODY_STATE)
        cons = np.ascontiguousarray(getattr(batch, "constraints", np.zeros(0, R.AVBD_CONSTRAINT)), R.AVBD_CONSTRAINT)
        offs = np.ascontiguousarray(batch.offsets, np.uint32)
        ns = len(offs) - 1
        if len(batch.hull):
            hull = np.ascontiguousarray(batch.hull, np.float32)
            self._check(self._lib.dio_world_upload_hulls(self._h, _lib.ptr(hull, ctypes.c_float), len(hull)))
        cts = (R.Container * ns)()
        for k, c in enumerate(batch.containers):
            if c is not None:
                cts[k] = c.to_c()
        self._check(
            self._lib.dio_batch_upload(self._h, _lib.ptr(bodies), _lib.ptr(state), len(bodies), _lib.ptr(cons),
                                       len(cons), _lib.ptr(offs, ctypes.c_uint32), ns, cts)
        )
        self._n = len(bodies)

    # ---- adjoint (frozen-contact replay of recorded steps)
    def adjoint_begin(self, max_steps: int, checkpoint_every: int = 8, max_recorded_contacts: int = 0) -> None:
        """Record the following steps for replay and reverse mode."""
        self._check(self._lib.dio_adjoint_begin(self._h, max_steps, checkpoint_every, max_recorded_contacts))

    def adjoint_end(self) -> None:
        self._check(self._lib.dio_adjoint_end(self._h))

    def adjoint_set_mode(self, mode: str) -> None:
        """What adjoint_backward differentiates: "exact" (the replay itself) or
        "frozen_multipliers" (the replay with the multipliers and penalties
        each dual update writes held as constants; bounded over long
        contact-rich runs where the exact derivative is not)."""
        code = {"exact": 0, "frozen_multipliers": 1}.get(mode)
        if code is None:
            raise ValueError(f"unknown adjoint mode {mode!r}: 'exact' or 'frozen_multipliers'")
        self._check(self._lib.dio_adjoint_set_mode(self._h, code))

    def adjoint_steps(self) -> int:
        v = ctypes.c_uint32()
        self._check(self._lib.dio_adjoint_steps(self._h, ctypes.byref(v)))
        return v.value

    def adjoint_replay(self, options: AvbdOptions, state0: np.ndarray | None = None):
        """Replay the recording (optionally from other velocities / masses):
        final positions (n x 3) and orientations (n x 4)."""
        n = self._n
        pos = np.zeros((n, 3), np.float32)
        rot = np.zeros((n, 4), np.float32)
        st = None if state0 is None else np.ascontiguousarray(state0, R.BODY_STATE)
        o = options.to_c()
        self._check(self._lib.dio_adjoint_replay(self._h, ctypes.byref(o), _lib.ptr(st),
                                                 _lib.ptr(pos, ctypes.c_float), _lib.ptr(rot, ctypes.c_float)))
        return pos, rot

    def adjoint_backward(self, options: AvbdOptions, d_positions=None, d_orientations=None, d_velocities=None,
                         d_angular=None, state0: np.ndarray | None = None) -> dict:
        """Reverse mode through the recording (the replay from state0, or the
        recorded run when None): adjoints of the initial state, the per-body
        inverse masses / moments and the friction, given the adjoint of the
        final state."""
        n = self._n
        st = None if state0 is None else np.ascontiguousarray(state0, R.BODY_STATE)
        f = lambda a, k: None if a is None else np.ascontiguousarray(a, np.float32).reshape(n, k)
        ins = [f(d_positions, 3), f(d_orientations, 4), f(d_velocities, 3), f(d_angular, 3)]
        out = {"positions": np.zeros((n, 3), np.float32), "orientations": np.zeros((n, 4), np.float32),
               "velocities": np.zeros((n, 3), np.float32), "angular_velocities": np.zeros((n, 3), np.float32),
               "prev_velocities": np.zeros((n, 3), np.float32), "inv_mass": np.zeros(n, np.float32),
               "inv_moment": np.zeros((n, 3), np.float32)}
        mu = ctypes.c_double()
        o = options.to_c()
        fp = lambda a: _lib.ptr(a, ctypes.c_float)
        self._check(self._lib.dio_adjoint_backward(
            self._h, ctypes.byref(o), _lib.ptr(st), *[fp(a) for a in ins], fp(out["positions"]), fp(out["orientations"]),
            fp(out["velocities"]), fp(out["angular_velocities"]), fp(out["prev_velocities"]), fp(out["inv_mass"]),
            fp(out["inv_moment"]), ctypes.byref(mu)))
        out["friction"] = mu.value
        return out

    def step(self, options: AvbdOptions, nsteps: int = 1, lag: int = 1, profile: bool = False,
             mode: int = 0) -> None:
        """dio_step: nsteps frames; mode 0 Jacobi (the reference), 1 coloured Gauss-Seidel."""
        so = R.StepOpts(1 if lag else 0, int(mode), 1 if profile else 0)
        o = options.to_c()
        self._check(self._lib.dio_step(self._h, ctypes.byref(o), ctypes.byref(so), nsteps))

    def set_tuning(self, primal_lanes: int = 0, broad_lanes: int = 0, resort_interval: int = 0,
                   fuse_dual: int | None = None, adjoint_lanes: int = 0, overlap_broad: int | None = None,
                   primal_order: int | None = None, scene_solve: int | None = None,
                   colour_lanes: int | None = None, primal_packed: int | None = None) -> None:
        """Launch shapes and schedule (0 / None keeps the current value).  A
        primal_lanes given without primal_packed selects that many lanes per
        body for the step's Jacobi primal (primal_packed 0)."""
        cur = self.tuning()
        if primal_lanes and primal_packed is None:
            primal_packed = 0
        t = R.Tuning()
        t.primal_lanes = primal_lanes or cur["primal_lanes"]
        t.broad_lanes = broad_lanes or cur["broad_lanes"]
        t.resort_interval = resort_interval or cur["resort_interval"]
        t.fuse_dual = cur["fuse_dual"] if fuse_dual is None else int(fuse_dual)
        t.adjoint_lanes = adjoint_lanes or cur["adjoint_lanes"]
        t.overlap_broad = cur["overlap_broad"] if overlap_broad is None else int(overlap_broad)
        t.primal_order = cur["primal_order"] if primal_order is None else int(primal_order)
        t.scene_solve = cur["scene_solve"] if scene_solve is None else int(scene_solve)
        t.colour_lanes = cur["colour_lanes"] if colour_lanes is None else int(colour_lanes)
        t.primal_packed = cur["primal_packed"] if primal_packed is None else int(primal_packed)
        self._check(self._lib.dio_set_tuning(self._h, ctypes.byref(t)))

    def tuning(self) -> dict:
        t = R.Tuning()
        self._check(self._lib.dio_get_tuning(self._h, ctypes.byref(t)))
        return {"primal_lanes": t.primal_lanes, "broad_lanes": t.broad_lanes, "resort_interval": t.resort_interval,
                "fuse_dual": t.fuse_dual, "adjoint_lanes": t.adjoint_lanes, "overlap_broad": t.overlap_broad,
                "primal_order": t.primal_order, "scene_solve": t.scene_solve, "colour_lanes": t.colour_lanes,
                "primal_packed": t.primal_packed}

    def autotune(self, options: AvbdOptions, reps: int = 20) -> dict:
        """Time every primal and broad-phase launch shape and both dual
        schedules on the current world, keep the fastest; returns the choice
        and the per-shape microseconds."""
        t = R.Tuning()
        us = (ctypes.c_float * 24)()
        o = options.to_c()
        self._check(self._lib.dio_autotune(self._h, ctypes.byref(o), reps, ctypes.byref(t), us))
        u = list(us)
        return {
            "primal_lanes": t.primal_lanes,
            "broad_lanes": t.broad_lanes,
            "fuse_dual": t.fuse_dual,
            "us_per_launch": dict(zip((1, 4, 8, 16, 32, 64), u[:6])),
            "broad_us_per_pass": dict(zip((1, 2, 4, 8, 16, 32), u[6:12])),
            "dual_schedule_us": {"unfused": u[12], "fused": u[13]},
            "primal_order": t.primal_order,
            "primal_order_us": {"index": u[14], "morton": u[15]},
            "colour_lanes": t.colour_lanes,
            "colour_us_per_iteration": dict(zip((1, 4, 8, 16, 32, 64), u[16:22])),
            "primal_packed": t.primal_packed,
            "packed_us": u[22],
            "span_us": u[23],
        }

    def stats(self) -> R.StepStats:
        s = R.StepStats()
        self._check(self._lib.dio_step_get_stats(self._h, ctypes.byref(s)))
        return s

    def download(self) -> dict:
        n = self._n
        out = {
            "positions": np.zeros((n, 3), np.float32),
            "orientations": np.zeros((n, 4), np.float32),
            "velocities": np.zeros((n, 3), np.float32),
            "angular_velocities": np.zeros((n, 3), np.float32),
            "prev_velocities": np.zeros((n, 3), np.float32),
        }
        fp = lambda a: _lib.ptr(a, ctypes.c_float)  # noqa: E731
        self._check(
            self._lib.dio_world_download(self._h, fp(out["positions"]), fp(out["orientations"]),
                                         fp(out["velocities"]), fp(out["angular_velocities"]),
                                         fp(out["prev_velocities"]))
        )
        return out

    def contacts(self) -> np.ndarray:
        m = ctypes.c_uint32()
        st = self._lib.dio_world_contacts(self._h, None, 0, ctypes.byref(m))
        if st not in (0, 1):
            self._check(st)
        out = np.zeros(m.value, R.AVBD_CONTACT)
        self._check(self._lib.dio_world_contacts(self._h, _lib.ptr(out), m.value, ctypes.byref(m)))
        return out

    def colours(self) -> tuple:
        """The last coloured step's colour per body (0xFFFFFFFE: static) and
        the number of colours."""
        out = np.zeros(self._n, np.uint32)
        nc = ctypes.c_uint32()
        self._check(self._lib.dio_world_colours(self._h, _lib.ptr(out, ctypes.c_uint32), self._n, ctypes.byref(nc)))
        return out, nc.value

    def joints(self) -> tuple:
        """Each constraint's multipliers after the last step: lambda and
        penalty per row, (c, 3) float32 each (zero for soft joints)."""
        c = ctypes.c_uint32()
        self._check(self._lib.dio_world_joints(self._h, None, None, 0, ctypes.byref(c)))
        lam = np.zeros((c.value, 3), np.float32)
        pen = np.zeros((c.value, 3), np.float32)
        self._check(self._lib.dio_world_joints(self._h, _lib.ptr(lam, ctypes.c_float), _lib.ptr(pen, ctypes.c_float),
                                               c.value, ctypes.byref(c)))
        return lam, pen

    def set_instances(self, scales, order=None) -> None:
        """The instanced renderer's data (engine/src/render/renderer.rs:222-255):
        per-body render scales (n x 3) and the instance order (order[s] is
        the body drawn in slot s; None: slot s draws body s, one slot per
        body).  Kept until the next call or upload."""
        sc = np.ascontiguousarray(scales, np.float32).reshape(-1, 3)
        if sc.shape[0] != self._n:
            raise ValueError(f"scales must hold 3 floats per body ({self._n} bodies)")
        if order is None:
            count, optr = self._n, None
        else:
            order = np.ascontiguousarray(order, np.uint32).reshape(-1)
            count, optr = order.size, _lib.ptr(order, ctypes.c_uint32)
        self._check(self._lib.dio_world_set_instances(self._h, _lib.ptr(sc, ctypes.c_float), optr, count))

    def instances(self, view_proj, dst: int, dst_bytes: int) -> None:
        """Enqueue, for every instance slot s, (view_proj @ model, model) into
        the device memory at address dst + 128 s: two 4 x 4 float matrices
        stored column by column (engine/shaders/mesh.vert's InstanceTransform),
        model built from the body's current pose and scale.  view_proj is the
        4 x 4 matrix projection @ view (rows by columns); dst is 16-byte
        aligned device memory of at least 128 bytes per slot."""
        vp = np.ascontiguousarray(np.asarray(view_proj, np.float32).reshape(4, 4).T).reshape(16)
        self._check(self._lib.dio_world_instances(self._h, _lib.ptr(vp, ctypes.c_float), ctypes.c_void_p(dst),
                                                  dst_bytes))

    def import_external(self, fd: int, object_size: int, offset: int = 0, size: int | None = None) -> tuple:
        """Map device memory another API exported as a file descriptor (a
        Vulkan buffer's opaque-fd memory, a DMA-BUF) into the context's device:
        returns (device address of bytes offset .. offset + size, handle).
        The library imports a duplicate; the caller keeps fd."""
        if size is None:
            size = object_size - offset
        ptr = ctypes.c_void_p()
        handle = ctypes.c_void_p()
        self._check(self._lib.dio_import_external_memory(self._h, fd, object_size, offset, size, ctypes.byref(ptr),
                                                         ctypes.byref(handle)))
        return ptr.value, handle

    def release_external(self, handle) -> None:
        self._check(self._lib.dio_release_external_memory(self._h, handle))

    def pairs(self) -> np.ndarray:
        n = ctypes.c_uint32()
        st = self._lib.dio_world_pairs(self._h, None, 0, ctypes.byref(n))
        if st not in (0, 1):
            self._check(st)
        out = np.zeros(n.value, R.GPU_PAIR)
        self._check(self._lib.dio_world_pairs(self._h, _lib.ptr(out), n.value, ctypes.byref(n)))
        return out


class AvbdSolver:
    """The reference's world object (physics/gpu.rs:158-312) stepped on the GPU.

    ``bodies``, ``container`` and ``constraints`` mirror the reference fields;
    ``step`` runs the demo frame (contact lag 1) on the device and syncs the
    non-static bodies back, as ``sync_bodies_from_gpu`` does."""

    def __init__(self, bodies: Sequence[AvbdBody], device: int = 0, **limits):
        self.bodies = list(bodies)
        self.container: Optional[AvbdContainer] = None
        self.constraints: list = []
        self._ctx: Optional[PhysicsContext] = None
        self._limits = limits
        self._device = device
        self._dirty = True

    def _scene(self, options: AvbdOptions) -> Scene:
        return Scene.from_bodies(self.bodies, self.constraints, self.container, options)

    def step(self, options: AvbdOptions, nsteps: int = 1) -> None:
        if self._ctx is None or self._ctx.max_bodies < len(self.bodies):
            self._ctx = PhysicsContext(max(len(self.bodies), 1), device=self._device, **self._limits)
            self._dirty = True
        if self._dirty:
            self._ctx.upload(self._scene(options))
            self._dirty = False
        self._ctx.step(options, nsteps)
        res = self._ctx.download()
        for i, b in enumerate(self.bodies):
            if b.is_static():
                continue
            b.position = tuple(map(float, res["positions"][i]))
            b.orientation = tuple(map(float, res["orientations"][i]))
            b.velocity = tuple(map(float, res["velocities"][i]))
            b.angular_velocity = tuple(map(float, res["angular_velocities"][i]))
            b.prev_velocity = tuple(map(float, res["prev_velocities"][i]))

    def contact_count(self) -> int:
        return 0 if self._ctx is None else len(self._ctx.contacts())

    def contacts(self) -> np.ndarray:
        return np.zeros(0, R.AVBD_CONTACT) if self._ctx is None else self._ctx.contacts()

    def mark_dirty(self) -> None:
        """Call after editing bodies / container / constraints on the host."""
        self._dirty = True


def quat_normalize(q: np.ndarray) -> np.ndarray:
    q = np.asarray(q, np.float64)
    return q / np.linalg.norm(q, axis=-1, keepdims=True)


def angle_of(q: np.ndarray) -> float:
    q = quat_normalize(q)
    return 2.0 * math.acos(min(1.0, abs(float(q[3]))))

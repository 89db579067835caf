"""Scene and trajectory files: a seeded world and its per-step poses (and,
optionally, velocities and broad-phase pairs) in one little-endian binary
file, readable from any language that knows the reference's records.

Layout (byte offsets; every record is the reference's #[repr(C)] layout, the
same bytes the C ABI takes):

    header, 64 bytes
      0  char[8]  "DIOTRAJ\\0"
      8  u32      version (1)
     12  u32      bodies n
     16  u32      constraints c
     20  u32      hull vertices h
     24  u32      scenes s (1 for a single world)
     28  u32      flags: 1 = frames carry velocities, 2 = frames carry pairs
     32  u32[5]   record sizes: GpuBody 112, GpuAvbdBodyState 80,
                  GpuAvbdConstraint 48, options 48, container 36
     52  u32[3]   reserved (0)
    scene
      options record (dt, gravity[3], iterations, alpha, beta_lin, margin,
        penalty_min, penalty_max, gamma, friction)
      GpuBody[n], GpuAvbdBodyState[n], GpuAvbdConstraint[c], f32[h][3]
      u32[s + 1] scene offsets (0 .. n), container[s] (enabled 0: none)
    frames, to the end of the file
      char[4] "FRAM", u32 pair count p (0 without flag 2), u64 step
      f32[n][3] positions, f32[n][4] orientations (x, y, z, w)
      flag 1: f32[n][3] velocities, f32[n][3] angular velocities
      flag 2: u32[p][2] pairs (a < b, sorted)

The writer appends frames as they come, so a long run need not fit in memory;
`read` loads the whole file.
"""

from __future__ import annotations

import dataclasses

import ctypes
import struct
from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional

import numpy as np

from . import records as R
from .physics import AvbdContainer, AvbdOptions, Scene

MAGIC = b"DIOTRAJ\x00"
VERSION = 1
FLAG_VELOCITIES = 1
FLAG_PAIRS = 2
_HEADER = struct.Struct("<8s6I5I3I")
_FRAME = struct.Struct("<4sIQ")
_SIZES = (R.GPU_BODY.itemsize, R.BODY_STATE.itemsize, R.AVBD_CONSTRAINT.itemsize, ctypes.sizeof(R.Options),
          ctypes.sizeof(R.Container))
assert _HEADER.size == 64


@dataclass
class Frame:
    step: int
    positions: np.ndarray             # (n, 3) float32
    orientations: np.ndarray          # (n, 4) float32
    velocities: Optional[np.ndarray] = None
    angular_velocities: Optional[np.ndarray] = None
    pairs: Optional[np.ndarray] = None  # (p, 2) uint32


@dataclass
class Trajectory:
    scene: Scene
    scene_offsets: np.ndarray                  # (s + 1,) uint32
    containers: list                           # s entries, AvbdContainer or None
    flags: int = 0
    frames: list = field(default_factory=list)


def _options_bytes(o: AvbdOptions) -> bytes:
    return bytes(o.to_c())


def _options_from(b: bytes) -> AvbdOptions:
    c = R.Options.from_buffer_copy(b)
    return AvbdOptions(dt=c.dt, gravity=tuple(c.gravity), iterations=c.iterations, alpha=c.alpha, beta_lin=c.beta_lin,
                       margin=c.margin, penalty_min=c.penalty_min, penalty_max=c.penalty_max, gamma=c.gamma,
                       friction=c.friction)


def _container_bytes(c: Optional[AvbdContainer]) -> bytes:
    return bytes(c.to_c()) if c is not None else bytes(R.Container())


def _container_from(b: bytes) -> Optional[AvbdContainer]:
    c = R.Container.from_buffer_copy(b)
    if not c.enabled:
        return None
    return AvbdContainer(center=tuple(c.center), rotation=tuple(c.rotation), half_extent=c.half_extent)


class Writer:
    """Write the header and scene at once, then frames one by one."""

    def __init__(self, path, scene: Scene, flags: int = 0, scene_offsets=None, containers=None):
        n = scene.n
        offs = np.array([0, n] if scene_offsets is None else scene_offsets, np.uint32)
        cts = [scene.container] if containers is None else list(containers)
        if offs[0] != 0 or offs[-1] != n or len(cts) != len(offs) - 1:
            raise ValueError("scene offsets must run from 0 to n with one container per scene")
        self.n, self.flags = n, int(flags)
        self._f = open(path, "wb")
        hull = np.ascontiguousarray(scene.hull, np.float32).reshape(-1, 3)
        cons = np.ascontiguousarray(scene.constraints, R.AVBD_CONSTRAINT)
        self._f.write(_HEADER.pack(MAGIC, VERSION, n, len(cons), len(hull), len(offs) - 1, self.flags, *_SIZES,
                                   0, 0, 0))
        self._f.write(_options_bytes(scene.options))
        self._f.write(np.ascontiguousarray(scene.bodies, R.GPU_BODY).tobytes())
        self._f.write(np.ascontiguousarray(scene.state, R.BODY_STATE).tobytes())
        self._f.write(cons.tobytes())
        self._f.write(hull.tobytes())
        self._f.write(offs.tobytes())
        for c in cts:
            self._f.write(_container_bytes(c))

    def frame(self, step: int, positions, orientations, velocities=None, angular_velocities=None, pairs=None):
        n = self.n
        pos = np.ascontiguousarray(positions, np.float32).reshape(n, 3)
        rot = np.ascontiguousarray(orientations, np.float32).reshape(n, 4)
        pr = None
        if self.flags & FLAG_PAIRS:
            if pairs is None:
                raise ValueError("this file's frames carry pairs")
            pr = np.ascontiguousarray(np.asarray(pairs).view(np.uint32) if np.asarray(pairs).dtype == R.GPU_PAIR
                                      else pairs, np.uint32).reshape(-1, 2)
        self._f.write(_FRAME.pack(b"FRAM", 0 if pr is None else len(pr), int(step)))
        self._f.write(pos.tobytes())
        self._f.write(rot.tobytes())
        if self.flags & FLAG_VELOCITIES:
            if velocities is None or angular_velocities is None:
                raise ValueError("this file's frames carry velocities")
            self._f.write(np.ascontiguousarray(velocities, np.float32).reshape(n, 3).tobytes())
            self._f.write(np.ascontiguousarray(angular_velocities, np.float32).reshape(n, 3).tobytes())
        if pr is not None:
            self._f.write(pr.tobytes())

    def close(self):
        if self._f:
            self._f.close()
            self._f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def write(path, traj: Trajectory) -> None:
    with Writer(path, traj.scene, traj.flags, traj.scene_offsets, traj.containers) as w:
        for f in traj.frames:
            w.frame(f.step, f.positions, f.orientations, f.velocities, f.angular_velocities, f.pairs)


def read(path) -> Trajectory:
    data = Path(path).read_bytes()
    if len(data) < _HEADER.size:
        raise ValueError("not a trajectory file: too short")
    magic, version, n, c, h, s, flags, *rest = _HEADER.unpack_from(data, 0)
    sizes = tuple(rest[:5])
    if magic != MAGIC:
        raise ValueError("not a trajectory file: bad magic")
    if version != VERSION:
        raise ValueError(f"unsupported trajectory version {version}")
    if sizes != _SIZES:
        raise ValueError(f"record sizes {sizes} differ from this build's {_SIZES}")
    if s < 1:
        raise ValueError("a trajectory holds at least one scene")
    off = _HEADER.size

    def take(count: int, dtype) -> np.ndarray:
        nonlocal off
        dt = np.dtype(dtype)
        nbytes = count * dt.itemsize
        if off + nbytes > len(data):
            raise ValueError("truncated trajectory file")
        a = np.frombuffer(data, dt, count, off).copy()
        off += nbytes
        return a

    opts = _options_from(take(_SIZES[3], np.uint8).tobytes())
    bodies = take(n, R.GPU_BODY)
    state = take(n, R.BODY_STATE)
    cons = take(c, R.AVBD_CONSTRAINT)
    hull = take(3 * h, np.float32).reshape(h, 3)
    offs = take(s + 1, np.uint32)
    cts = [_container_from(take(_SIZES[4], np.uint8).tobytes()) for _ in range(s)]
    if offs[0] != 0 or offs[-1] != n or np.any(np.diff(offs.astype(np.int64)) < 0):
        raise ValueError("bad scene offsets")
    scene = Scene(bodies, state, cons, cts[0] if s == 1 else None, hull, opts, name=str(Path(path).stem))
    traj = Trajectory(scene, offs, cts, flags)
    while off < len(data):
        if off + _FRAME.size > len(data):
            raise ValueError("truncated trajectory file")
        tag, p, step = _FRAME.unpack_from(data, off)
        off += _FRAME.size
        if tag != b"FRAM":
            raise ValueError("bad frame tag")
        fr = Frame(int(step), take(3 * n, np.float32).reshape(n, 3), take(4 * n, np.float32).reshape(n, 4))
        if flags & FLAG_VELOCITIES:
            fr.velocities = take(3 * n, np.float32).reshape(n, 3)
            fr.angular_velocities = take(3 * n, np.float32).reshape(n, 3)
        if flags & FLAG_PAIRS:
            fr.pairs = take(2 * p, np.uint32).reshape(p, 2)
        traj.frames.append(fr)
    return traj


def record(ctx, scene: Scene, path, steps: int, every: int = 1, velocities: bool = False, pairs: bool = False,
           options: Optional[AvbdOptions] = None) -> int:
    """Upload scene to a PhysicsContext, step it and write a frame after every
    `every` steps (the initial state is frame 0).  The file's header names the
    options the frames were stepped with: `options` when given, else the
    scene's own.  Returns the frames written."""
    flags = (FLAG_VELOCITIES if velocities else 0) | (FLAG_PAIRS if pairs else 0)
    opts = options or scene.options
    ctx.upload(scene)
    frames = 0
    with Writer(path, dataclasses.replace(scene, options=opts), flags) as w:
        def put(step: int) -> None:
            d = ctx.download()
            pr = None
            if pairs:
                pr = ctx.pairs() if step > 0 else np.zeros(0, R.GPU_PAIR)
            w.frame(step, d["positions"], d["orientations"], d.get("velocities"), d.get("angular_velocities"), pr)

        put(0)
        frames = 1
        done = 0
        while done < steps:
            k = min(every, steps - done)
            ctx.step(opts, k)
            done += k
            put(done)
            frames += 1
    return frames

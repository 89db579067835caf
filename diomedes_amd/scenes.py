"""Seeded scene generators for the benchmark and parity workloads.

Randomness is SplitMix64 (Steele, Lea & Flood 2014): draw i of a stream seeded
with s is mix(s + (i + 1) * 0x9E3779B97F4A7C15) with the published
finaliser, and a uniform in [0, 1) is (draw >> 11) * 2**-53.  Draws are
consumed in the documented order of each generator, computed in float64 and
rounded once to float32 (the ABI's type).  Random orientations are
Shoemake's uniform quaternions (Graphics Gems III, 1992).

Workloads (all bodies dynamic unless stated; ground = the floor y = 0 of an
AvbdContainer whose other walls are out of reach):

- ``box_stack``: 10 unit cubes (h 0.5, mass 1) stacked at y = 0.5 + i with
  U(-1e-3, 1e-3) jitter per coordinate; g = -9.81, dt = 1/60, 10 iterations,
  friction 0.5.
- ``cube_field``: 32 x 32 x 8 = 8192 cubes, h ~ U(0.25, 0.75), density 1
  (mass (2h)^3), grid spacing 2.0 with U(-0.1, 0.1) jitter, lowest layer at
  y = 2.0, uniform random orientations; g = -9.81, dt = 1/60, 4 iterations,
  friction 0.6.
"""

from __future__ import annotations

import math

import numpy as np

from . import records as R
from .physics import AvbdContainer, AvbdOptions, Scene

GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """Counter-form SplitMix64 stream with vectorised draws."""

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0

    def u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            i = np.arange(self.count + 1, self.count + n + 1, dtype=np.uint64)
            z = self.seed + i * GAMMA
            z = (z ^ (z >> np.uint64(30))) * M1
            z = (z ^ (z >> np.uint64(27))) * M2
            z = z ^ (z >> np.uint64(31))
        self.count += n
        return z

    def uniform(self, lo: float, hi: float, n: int) -> np.ndarray:
        u = (self.u64(n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
        return lo + (hi - lo) * u

    def normal(self, mean: float, std: float, n: int) -> np.ndarray:
        """Box-Muller pairs."""
        m = (n + 1) // 2
        u1 = self.uniform(0.0, 1.0, m)
        u2 = self.uniform(0.0, 1.0, m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        z = np.concatenate([r * np.cos(2 * math.pi * u2), r * np.sin(2 * math.pi * u2)])[:n]
        return mean + std * z


def shoemake(rng: SplitMix64, n: int) -> np.ndarray:
    """Uniform random unit quaternions (x, y, z, w)."""
    u1, u2, u3 = rng.uniform(0, 1, n), rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    q = np.stack(
        [a * np.sin(2 * math.pi * u2), a * np.cos(2 * math.pi * u2), b * np.sin(2 * math.pi * u3),
         b * np.cos(2 * math.pi * u3)],
        axis=1,
    )
    return q.astype(np.float32)


def _cube_inv_moment(mass: np.ndarray, h: np.ndarray) -> np.ndarray:
    """1 / ((2/3) m h^2) in float32, in AvbdBody::cube's order (physics/gpu.rs:55)."""
    m = mass.astype(np.float32)
    hh = h.astype(np.float32)
    return (np.float32(1.0) / (((np.float32(2.0 / 3.0) * m) * hh) * hh)).astype(np.float32)


def cubes(positions, half, mass, orientations=None) -> tuple[np.ndarray, np.ndarray]:
    """GpuBody + GpuAvbdBodyState records of dynamic cubes."""
    n = len(positions)
    bodies = np.zeros(n, R.GPU_BODY)
    bodies["position"][:, :3] = np.asarray(positions, np.float32)
    if orientations is None:
        bodies["orientation"][:, 3] = 1.0
    else:
        bodies["orientation"] = np.asarray(orientations, np.float32)
    bodies["tag"] = R.CUBE
    h = np.asarray(half, np.float32)
    bodies["corners"][:, 0, :3] = h[:, None]
    state = np.zeros(n, R.BODY_STATE)
    m = np.asarray(mass, np.float32)
    state["inv_mass"] = np.float32(1.0) / m
    state["inv_moment"][:, :3] = _cube_inv_moment(m, h)[:, None]
    state["friction"] = np.float32(0.8)
    return bodies, state


def floor_container(extent: float) -> AvbdContainer:
    """A container whose floor is y = 0 and whose walls are `extent` away."""
    return AvbdContainer(center=(0.0, float(extent), 0.0), rotation=(0.0, 0.0, 0.0, 1.0), half_extent=float(extent))


def box_stack(seed: int = 0, count: int = 10) -> Scene:
    rng = SplitMix64(seed)
    jitter = rng.uniform(-1e-3, 1e-3, 3 * count).reshape(count, 3)
    pos = np.stack([np.zeros(count), 0.5 + np.arange(count), np.zeros(count)], axis=1) + jitter
    bodies, state = cubes(pos, np.full(count, 0.5), np.ones(count))
    opts = AvbdOptions(dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), iterations=10, friction=0.5)
    return Scene(bodies, state, container=floor_container(50.0), options=opts, name=f"box_stack{count}_s{seed}")


def cube_field(seed: int = 0, nx: int = 32, ny: int = 8, nz: int = 32, spacing: float = 2.0) -> Scene:
    """Draw order: half extents (n), jitter (3n, x y z per body), orientations
    (3n Shoemake uniforms); bodies in x-major, then z, then layer order."""
    rng = SplitMix64(seed)
    n = nx * ny * nz
    half = rng.uniform(0.25, 0.75, n)
    jitter = rng.uniform(-0.1, 0.1, 3 * n).reshape(n, 3)
    q = shoemake(rng, n)
    iy, iz, ix = np.meshgrid(np.arange(ny), np.arange(nz), np.arange(nx), indexing="ij")
    base = np.stack(
        [(ix.ravel() - (nx - 1) / 2.0) * spacing, 2.0 + iy.ravel() * spacing, (iz.ravel() - (nz - 1) / 2.0) * spacing],
        axis=1,
    )
    pos = base + jitter
    h32 = half.astype(np.float32)
    mass = (np.float32(2.0) * h32) ** 3
    bodies, state = cubes(pos, h32, mass, q)
    extent = max(nx, nz) * spacing + 8.0
    opts = AvbdOptions(dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), iterations=4, friction=0.6)
    return Scene(bodies, state, container=floor_container(extent), options=opts,
                 name=f"cube_field{nx}x{nz}x{ny}_s{seed}")

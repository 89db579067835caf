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
- ``mixed_field``: n bodies, a third each of subdivision-1 icosphere hulls
  (42 vertices, radius U(0.3, 0.6)), regular tetrahedra (edge U(0.5, 1.0))
  and boxes (half extents U(0.2, 0.5)), density 1, velocities N(0, 2) per
  axis, on a grid of spacing 2.0 inside a 60 m container (centre (0, 30, 0),
  half 30); plus ropes of rope_links small spheres (r 0.05, mass 0.1, the
  links are bodies) joined by Rope(max 0.2) constraints, and Hookean springs
  (k 1e4, rest = initial distance) between random grid neighbours;
  dt = 1/120, 8 iterations.
- ``batched_scenes``: a scene table of independent drops; scene sizes
  round(exp(U(ln 8, ln 512))); each scene's unit cubes (h 0.5, mass 1) and
  subdivision-1 icosphere hulls (r 0.5, density 1) fall on its own floor
  (see ``BatchedScenes``).
- ``wall_throw``: a 64 x 32 wall of bricks (0.5 x 0.25 x 0.25, density 1) on
  the floor, hit by a sphere (r 0.5, mass 20) thrown with v0 ~ (15 +- 2,
  2 +- 0.5, 0); dt = 1/60, 6 iterations, the reference demo's other options
  (beta 5000, margin 0.002, penalty_max 1e6, gamma 0.995, friction 0.25,
  g = -10).
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


# ------------------------------------------------------------- mixed convex
def icosphere(subdivisions: int = 1) -> np.ndarray:
    """Unit icosphere vertices: the icosahedron, each edge split at its
    midpoint `subdivisions` times, projected to the sphere (42 for 1)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = [np.array(x, np.float64) / np.linalg.norm(x) for x in v]
    for _ in range(subdivisions):
        cache, nf = {}, []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(verts)


def regular_tetrahedron(edge: float) -> np.ndarray:
    c = np.array([(1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)], np.float64)
    return c * (edge / (2.0 * 2 ** 0.5))


def mixed_field(seed: int = 0, n: int = 20000, ropes: int = 128, rope_links: int = 64, springs: int = 512,
                extent: float = 30.0) -> Scene:
    """Draw order: shape parameters (n), orientations (3n), velocities (3n
    normals), spring endpoints (2 x springs)."""
    rng = SplitMix64(seed)
    kinds = np.arange(n) % 3  # 0 hull, 1 tetrahedron, 2 box
    size = rng.uniform(0.0, 1.0, n)
    q = shoemake(rng, n)
    vel = rng.normal(0.0, 2.0, 3 * n).reshape(n, 3)
    side = int(np.ceil(n ** (1.0 / 3.0)))
    idx = np.arange(n)
    spacing = min(2.0, 1.8 * extent / side)
    base = np.stack([idx % side, (idx // side) % side, idx // (side * side)], axis=1).astype(np.float64)
    pos = (base - (side - 1) / 2.0) * spacing
    pos[:, 1] += extent
    ico = icosphere(1).astype(np.float32)
    n_rope_bodies = ropes * rope_links
    total = n + n_rope_bodies
    bodies = np.zeros(total, R.GPU_BODY)
    state = np.zeros(total, R.BODY_STATE)
    hull_pool = []
    for i in range(n):
        k = kinds[i]
        if k == 0:
            # hull vertices are not scaled per body: each hull body has its own 42 vertices in the pool
            r = np.float32(0.3 + 0.3 * size[i])
            verts = (ico * r).astype(np.float32)
            bodies[i]["tag"] = R.HULL
            bodies[i]["corners"][0, 0] = np.float32(len(hull_pool) * 42)
            bodies[i]["corners"][0, 1] = np.float32(42)
            hull_pool.append(verts)
            mass = np.float32(4.0 / 3.0 * np.pi) * r * r * r
            inv_i = np.float32(1.0) / (np.float32(0.4) * mass * r * r)
            state[i]["inv_moment"][:3] = inv_i
        elif k == 1:
            e = 0.5 + 0.5 * size[i]
            c = regular_tetrahedron(e).astype(np.float32)
            bodies[i]["tag"] = R.TETRAHEDRON
            bodies[i]["corners"][:, :3] = c
            mass = np.float32(e ** 3 / (6.0 * 2 ** 0.5))
            radius = np.float32(np.max(np.sqrt(np.sum(c * c, axis=1))))
            state[i]["inv_moment"][:3] = np.float32(1.0) / (np.float32(0.4) * mass * radius * radius)
        else:
            h = np.float32([0.2 + 0.3 * size[i], 0.2 + 0.3 * ((size[i] * 7.31) % 1.0), 0.2 + 0.3 * ((size[i] * 3.17) % 1.0)])
            bodies[i]["tag"] = R.BOX
            bodies[i]["corners"][0, :3] = h
            mass = np.float32(8.0) * h[0] * h[1] * h[2]
            for ax, (j, k2) in enumerate(((1, 2), (0, 2), (0, 1))):
                state[i]["inv_moment"][ax] = np.float32(1.0) / (mass / np.float32(3.0) * (h[j] * h[j] + h[k2] * h[k2]))
        state[i]["inv_mass"] = np.float32(1.0) / np.float32(mass)
    bodies["position"][:n, :3] = pos.astype(np.float32)
    bodies["orientation"][:n] = q
    state["vel"][:n, :3] = vel.astype(np.float32)
    state["friction"] = np.float32(0.8)
    # ropes: chains of small spheres hanging along -y from the top of the volume
    cons = []
    rb = n
    for r_i in range(ropes):
        x0 = -extent * 0.9 + (2 * extent * 0.9) * (r_i % 16) / 15.0
        z0 = -extent * 0.9 + (2 * extent * 0.9) * (r_i // 16) / max(1, (ropes - 1) // 16)
        for l in range(rope_links):
            b = rb + r_i * rope_links + l
            bodies[b]["tag"] = R.SPHERE
            bodies[b]["corners"][0, 0] = np.float32(0.05)
            bodies[b]["position"][:3] = (x0, 2 * extent - 1.0 - 0.2 * l, z0)
            bodies[b]["orientation"][3] = 1.0
            m = np.float32(0.1)
            state[b]["inv_mass"] = np.float32(1.0) / m
            state[b]["inv_moment"][:3] = np.float32(1.0) / (np.float32(0.4) * m * np.float32(0.05) * np.float32(0.05))
            if l > 0:
                cons.append((b - 1, b, R.ROPE, 0.2, 1000.0))
    # springs between grid neighbours (index + 1 along x)
    a_idx = rng.uniform(0, 1, springs)
    for s_i in range(springs):
        a = int(a_idx[s_i] * (n - 1))
        if (a + 1) % side == 0:
            a -= 1
        b = a + 1
        rest = float(np.linalg.norm(pos[b] - pos[a]))
        cons.append((a, b, R.SPRING, rest, 1.0e4))
    cons_arr = np.zeros(len(cons), R.AVBD_CONSTRAINT)
    for i, (a, b, kind, length, k) in enumerate(cons):
        cons_arr[i]["a"], cons_arr[i]["b"], cons_arr[i]["kind"] = a, b, kind
        cons_arr[i]["anchor_a"][3] = length
        cons_arr[i]["anchor_b"][3] = k
    hull = np.concatenate(hull_pool).astype(np.float32) if hull_pool else np.zeros((0, 3), np.float32)
    container = AvbdContainer(center=(0.0, extent, 0.0), rotation=(0.0, 0.0, 0.0, 1.0), half_extent=extent)
    opts = AvbdOptions(dt=1.0 / 120.0, gravity=(0.0, -9.81, 0.0), iterations=8)
    return Scene(bodies, state, cons_arr, container, hull, opts, name=f"mixed_field{n}_s{seed}")


# ------------------------------------------------------------- hard chains
def hard_chains(seed: int = 0, chains: int = 256, links: int = 32, half: float = 0.15, gap: float = 0.04,
                row_pitch: float = 1.0, bend: str = "", bend_stiffness: float = 5.0) -> Scene:
    """Chains of cubes on hard joints, each hung from a static anchor and
    released horizontally: they swing down in parallel planes `row_pitch`
    apart, neighbouring rows in opposite directions, and their lower links
    pile on the floor into one another.  No two bodies overlap at the start:
    anchors sit in rows (z) of up to 8 chains whose spans (x) do not overlap,
    at a height of 0.8 of a chain's length.  Chain k uses hard sockets
    between link ends (k % 3 == 0), hard rods (1) or hard ropes (2) between
    link centres.  bend "soft" or "hard" adds a bending joint beside each hard
    socket (an angular spring of bend_stiffness, or multipliers), holding each
    link's orientation to its predecessor's.  Bodies are the anchors (0 ..
    chains - 1), then the links chain by chain.  Draw order: jitter (3 x
    chains x links).  Steps need the coloured mode (hard joints join dynamic
    bodies)."""
    rng = SplitMix64(seed)
    jitter = rng.uniform(-1e-3, 1e-3, 3 * chains * links).reshape(chains * links, 3)
    step = 2.0 * half + gap
    span = links * step
    height = 0.8 * span
    cols = min(8, chains)
    rows = (chains + cols - 1) // cols
    n = chains + chains * links
    bodies = np.zeros(n, R.GPU_BODY)
    state = np.zeros(n, R.BODY_STATE)
    bodies["orientation"][:, 3] = 1.0
    anchors = np.zeros((chains, 3))
    dirs = np.zeros((chains, 3))
    for k in range(chains):
        r, c = k // cols, k % cols
        # even rows point +x from the span's left end, odd rows -x from its right end
        sgn = 1.0 if r % 2 == 0 else -1.0
        x0 = (c - (cols - 1) / 2.0) * (span + 1.0) - sgn * span / 2.0
        anchors[k] = (x0, height, (r - (rows - 1) / 2.0) * row_pitch)
        dirs[k] = (sgn, 0.0, 0.0)
    bodies["position"][:chains, :3] = anchors.astype(np.float32)
    bodies["tag"][:chains] = R.SPHERE
    bodies["corners"][:chains, 0, 0] = np.float32(0.1)
    link_pos = np.zeros((chains * links, 3))
    for k in range(chains):
        for i in range(links):
            link_pos[k * links + i] = anchors[k] + dirs[k] * step * (i + 1)
    link_pos += jitter
    lb, ls = cubes(link_pos, np.full(chains * links, half), np.ones(chains * links))
    bodies[chains:] = lb
    state[chains:] = ls
    state["friction"] = np.float32(0.6)
    if bend not in ("", "soft", "hard"):
        raise ValueError("bend must be '', 'soft' or 'hard'")
    n_bend = links * ((chains + 2) // 3) if bend else 0
    cons = np.zeros(chains * links + n_bend, R.AVBD_CONSTRAINT)
    c = 0
    for k in range(chains):
        d = dirs[k].astype(np.float32)
        kind = (R.HARD_SOCKET, R.HARD_ROD, R.HARD_ROPE)[k % 3]
        for i in range(links):
            a = k if i == 0 else chains + k * links + i - 1
            b = chains + k * links + i
            cons[c]["a"], cons[c]["b"], cons[c]["kind"] = a, b, kind
            if kind == R.HARD_SOCKET:
                # this link's near end to the previous link's far end (or the
                # point half a link from the anchor)
                cons[c]["anchor_a"][:3] = d * np.float32(step / 2.0)
                cons[c]["anchor_b"][:3] = -d * np.float32(step / 2.0)
            else:
                cons[c]["anchor_a"][3] = np.float32(step)
            c += 1
            if bend and kind == R.HARD_SOCKET:
                # every body starts unrotated: the rest rotation is the identity
                cons[c]["a"], cons[c]["b"] = a, b
                cons[c]["kind"] = R.BEND if bend == "soft" else R.HARD_BEND
                cons[c]["anchor_a"][3] = 1.0
                cons[c]["anchor_b"][3] = np.float32(bend_stiffness) if bend == "soft" else 0.0
                c += 1
    assert c == len(cons)
    extent = max(cols * (span + 1.0), rows * row_pitch) / 2.0 + span + 4.0
    opts = AvbdOptions(dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), iterations=8, friction=0.6)
    return Scene(bodies, state, cons, floor_container(extent), options=opts,
                 name=f"hard_chains{chains}x{links}{('_' + bend + '_bend') if bend else ''}_s{seed}")


# ------------------------------------------------------------- wall throw
def wall_throw(seed: int = 0, columns: int = 64, rows: int = 32) -> Scene:
    """Bricks 0.5 (z) x 0.25 (y) x 0.25 (x) stacked in a wall in the y-z
    plane at x = 0; the sphere starts at (-4, 1.5, 0).  Draw order: v0.x,
    v0.y."""
    rng = SplitMix64(seed)
    vx = rng.uniform(13.0, 17.0, 1)[0]
    vy = rng.uniform(1.5, 2.5, 1)[0]
    h = np.float32([0.125, 0.125, 0.25])
    n = columns * rows
    bodies = np.zeros(n + 1, R.GPU_BODY)
    state = np.zeros(n + 1, R.BODY_STATE)
    r_, c_ = np.meshgrid(np.arange(rows), np.arange(columns), indexing="ij")
    bodies["position"][:n, 0] = 0.0
    bodies["position"][:n, 1] = (h[1] + 2 * h[1] * r_.ravel()).astype(np.float32)
    bodies["position"][:n, 2] = ((c_.ravel() - (columns - 1) / 2.0) * 2 * h[2]).astype(np.float32)
    bodies["orientation"][:, 3] = 1.0
    bodies["tag"][:n] = R.BOX
    bodies["corners"][:n, 0, :3] = h
    mass = np.float32(8.0) * h[0] * h[1] * h[2]
    state["inv_mass"][:n] = np.float32(1.0) / mass
    for ax, (j, k) in enumerate(((1, 2), (0, 2), (0, 1))):
        state["inv_moment"][:n, ax] = np.float32(1.0) / (mass / np.float32(3.0) * (h[j] * h[j] + h[k] * h[k]))
    # projectile
    r, m = np.float32(0.5), np.float32(20.0)
    bodies[n]["tag"] = R.SPHERE
    bodies[n]["corners"][0, 0] = r
    bodies[n]["position"][:3] = (-4.0, 1.5, 0.0)
    state[n]["inv_mass"] = np.float32(1.0) / m
    state[n]["inv_moment"][:3] = np.float32(1.0) / (np.float32(0.4) * m * r * r)
    state[n]["vel"][:3] = (vx, vy, 0.0)
    state["friction"] = np.float32(0.8)
    # the reference demo's preset (demos/wall_throw/src/main.rs:61-68) with 6 iterations
    opts = AvbdOptions(dt=1.0 / 60.0, gravity=(0.0, -10.0, 0.0), iterations=6, alpha=0.99, beta_lin=5000.0,
                       margin=0.002, penalty_max=1.0e6, gamma=0.995, friction=0.25)
    return Scene(bodies, state, container=floor_container(40.0), options=opts, name=f"wall_throw{n}_s{seed}")


# ------------------------------------------------------------- batched scenes
class BatchedScenes:
    """A scene table: scene s owns bodies offsets[s] .. offsets[s+1] - 1 of the
    concatenated arrays and has its own container (floor).  All scenes share
    one hull vertex pool (the r = 0.5 icosphere) and one set of options."""

    def __init__(self, scenes: list, hull: np.ndarray, options: AvbdOptions):
        self.scenes = scenes
        self.hull = hull
        self.options = options
        self.offsets = np.zeros(len(scenes) + 1, np.uint32)
        self.offsets[1:] = np.cumsum([sc.n for sc in scenes])
        self.bodies = np.concatenate([sc.bodies for sc in scenes])
        self.state = np.concatenate([sc.state for sc in scenes])
        self.containers = [sc.container for sc in scenes]

    @property
    def n(self) -> int:
        return int(self.offsets[-1])

    @property
    def n_dynamic(self) -> int:
        return int(np.count_nonzero(self.state["inv_mass"] != 0))

    def scene_slice(self, s: int) -> slice:
        return slice(int(self.offsets[s]), int(self.offsets[s + 1]))


def batched_scenes(seed: int = 0, num_scenes: int = 4096, n_min: int = 8, n_max: int = 512) -> BatchedScenes:
    """Draw order per scene: the body count (1), kinds (n), position jitter
    (3n), orientations (Shoemake, 3n).  Bodies sit on a square grid of
    spacing 1.5 (side ceil(sqrt(n / 2)), so about two layers), the lowest
    layer at y = 1.0, with U(-0.1, 0.1) jitter; the container's floor is y = 0
    and its walls stand 3 m beyond the grid.  g = -9.81, dt = 1/60, 4
    iterations, friction 0.5."""
    rng = SplitMix64(seed)
    ico = (icosphere(1) * 0.5).astype(np.float32)
    r = np.float32(0.5)
    h = np.float32(0.5)
    m_hull = np.float32(4.0 / 3.0 * np.pi) * r * r * r
    inv_i_hull = np.float32(1.0) / (np.float32(0.4) * m_hull * r * r)
    m_cube = np.float32(1.0)
    inv_i_cube = np.float32(1.0) / (np.float32(2.0 / 3.0) * m_cube * h * h)
    opts = AvbdOptions(dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), iterations=4, friction=0.5)
    scenes = []
    for s in range(num_scenes):
        u = rng.uniform(0.0, 1.0, 1)[0]
        n = int(round(math.exp(math.log(n_min) + u * (math.log(n_max) - math.log(n_min)))))
        kinds = rng.uniform(0.0, 1.0, n) < 0.5  # True: icosphere hull
        jit = rng.uniform(-0.1, 0.1, 3 * n).reshape(n, 3)
        q = shoemake(rng, n)
        side = int(math.ceil(math.sqrt(n / 2.0)))
        idx = np.arange(n)
        gx, gz, gy = idx % side, (idx // side) % side, idx // (side * side)
        pos = np.stack([(gx - (side - 1) / 2.0) * 1.5, 1.0 + gy * 1.5, (gz - (side - 1) / 2.0) * 1.5], axis=1) + jit
        bodies = np.zeros(n, R.GPU_BODY)
        state = np.zeros(n, R.BODY_STATE)
        bodies["position"][:, :3] = pos.astype(np.float32)
        bodies["orientation"] = q
        bodies["tag"] = np.where(kinds, R.HULL, R.CUBE)
        bodies["corners"][kinds, 0, 0] = np.float32(0.0)
        bodies["corners"][kinds, 0, 1] = np.float32(len(ico))
        bodies["corners"][~kinds, 0, :3] = h
        state["inv_mass"] = np.where(kinds, np.float32(1.0) / m_hull, np.float32(1.0) / m_cube)
        state["inv_moment"][:, :3] = np.where(kinds, inv_i_hull, inv_i_cube)[:, None]
        state["friction"] = np.float32(0.5)
        half = float(max(8.0, (side - 1) * 0.75 + 3.0))
        cont = AvbdContainer(center=(0.0, half, 0.0), rotation=(0.0, 0.0, 0.0, 1.0), half_extent=half)
        scenes.append(Scene(bodies, state, container=cont, hull=ico, options=opts, name=f"batch{s}_n{n}_s{seed}"))
    return BatchedScenes(scenes, ico, opts)


def wall_throw_target(scene: Scene, seed: int) -> np.ndarray:
    """The seeded target field of the differentiable wall throw: each brick's
    initial centre plus U(-0.5, 0.5) per axis (draws after the scene's own,
    from stream seed + 1), shape (bricks, 3), float64."""
    rng = SplitMix64(seed + 1)
    nb = scene.n - 1
    return scene.bodies["position"][:nb, :3].astype(np.float64) + rng.uniform(-0.5, 0.5, 3 * nb).reshape(nb, 3)

"""dio_narrow_pairs_noalloc against dio_narrow_submit_pairs + dio_narrow_read:
the allocation-free narrow phase gives the same slots, bit for bit, through
the C ABI, on two calls in a row (the second with fewer bodies and pairs, so
it runs on the array the first call left behind), and its limits and input
checks answer before any launch."""

from __future__ import annotations

import numpy as np
import pytest

from diomedes_amd import PhysicsContext, scenes
from diomedes_amd import _lib
from diomedes_amd import records as R
from diomedes_amd.physics import AvbdBody, Scene, Shape, gpu_bodies

pytestmark = pytest.mark.gpu

MAX_BODIES, MAX_PAIRS, MAX_HULL = 16384, 65536, 16384
STATUS = {0: "OK", 1: "CAPACITY", 2: "INVALID", 3: "HIP", 4: "NOT_READY"}


def _fn():
    return _lib.load().dio_narrow_pairs_noalloc


def _noalloc(bodies, pairs, hull, cap=None):
    bodies = np.ascontiguousarray(bodies, R.GPU_BODY)
    pairs = np.ascontiguousarray(pairs, R.GPU_PAIR)
    hull = np.ascontiguousarray(hull, np.float32).reshape(-1, 3)
    cap = 4 * len(pairs) if cap is None else cap
    out = np.zeros(max(cap, 1), R.GPU_CONTACT_OUT)
    st = _fn()(0, bodies.ctypes.data, len(bodies), pairs.ctypes.data, len(pairs), hull.ctypes.data, len(hull),
               out.ctypes.data, cap)
    return STATUS[st], out[:4 * len(pairs)]


def _reference(bodies, pairs, hull):
    """The context's entry points: dio_narrow_submit_pairs + dio_narrow_read."""
    n = len(bodies)
    with PhysicsContext(n, max_pairs=max(len(pairs), 1), max_hull_vertices=max(len(hull), 1)) as ctx:
        if len(hull):
            ctx.upload(Scene(bodies, np.zeros(n, R.BODY_STATE), hull=hull))  # the pool the granular path reads
        return ctx.narrow_phase_run(bodies, pairs)


def _workload(seed, count):
    """Overlapping pairs of every class: box-box (SAT), sphere-sphere, and
    tetrahedra, icosphere hulls, spheres and boxes mixed (GJK + EPA), plus
    deeply overlapping 642-vertex hulls that take the second EPA pass."""
    rng = np.random.default_rng(seed)
    ico1 = (scenes.icosphere(1) * 0.5).astype(np.float32)
    ico3 = (scenes.icosphere(3) * 0.6).astype(np.float32)
    hull = np.concatenate([ico1, ico3])

    def shape(kind):
        if kind == 0:
            c = rng.normal(size=(4, 3)) * 0.5
            return Shape.tetrahedron(c - c.mean(axis=0))
        if kind == 1:
            return Shape.hull(0, len(ico1))
        if kind == 2:
            return Shape.sphere(rng.uniform(0.2, 0.6))
        if kind == 3:
            return Shape.box(rng.uniform(0.2, 0.5, 3))
        return Shape.hull(len(ico1), len(ico3))

    bodies = []
    for k in range(count):
        ka, kb = rng.integers(0, 5, 2)
        q = rng.normal(size=(2, 4))
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        pa = rng.normal(size=3) * 4.0
        scale = 0.2 if (ka == 4 and kb == 4) else 0.4
        bodies.append(AvbdBody(shape(ka), tuple(pa), tuple(q[0])))
        bodies.append(AvbdBody(shape(kb), tuple(pa + rng.normal(size=3) * scale), tuple(q[1])))
    bodies = gpu_bodies(bodies)
    n = len(bodies)
    pairs = np.zeros(n // 2, R.GPU_PAIR)
    pairs["a"] = np.arange(0, n, 2)
    pairs["b"] = np.arange(1, n, 2)
    # a few pairs across the list, and repeated pairs
    extra = np.zeros(64, R.GPU_PAIR)
    extra["a"] = rng.integers(0, n, 64)
    extra["b"] = rng.integers(0, n, 64)
    return bodies, np.concatenate([pairs, extra]), hull


def test_noalloc_equals_the_context_path_on_two_calls():
    for seed, count in ((1, 900), (2, 150)):
        bodies, pairs, hull = _workload(seed, count)
        print("noalloc call:", len(bodies), "bodies,", len(pairs), "pairs", flush=True)
        st, got = _noalloc(bodies, pairs, hull)
        assert st == "OK"
        want = _reference(bodies, pairs, hull)
        assert got.tobytes() == want.tobytes()
        assert int(np.sum(got["valid"])) > len(pairs) // 2


def test_noalloc_on_a_settled_cube_field():
    """Box manifolds of a settled field's pair list, then a smaller call."""
    from oracle import pyoracle

    pyoracle.load()
    for nx, steps in ((10, 60), (4, 30)):
        scene = scenes.cube_field(3, nx=nx, ny=3, nz=nx)
        w = pyoracle.OracleWorld(scene, 32)
        w.step(nsteps=steps)
        st = w.state()
        bodies = scene.bodies.copy()
        bodies["position"][:, :3] = st["positions"].astype(np.float32)
        bodies["orientation"] = st["orientations"].astype(np.float32)
        ref_pairs = pyoracle.broad_pairs(bodies)
        pairs = np.zeros(len(ref_pairs), R.GPU_PAIR)
        pairs["a"], pairs["b"] = ref_pairs[:, 0], ref_pairs[:, 1]
        stt, got = _noalloc(bodies, pairs, np.zeros((0, 3), np.float32))
        assert stt == "OK"
        assert got.tobytes() == _reference(bodies, pairs, np.zeros((0, 3), np.float32)).tobytes()


def test_noalloc_limits_and_checks():
    bodies, pairs, hull = _workload(3, 20)
    assert _noalloc(bodies, pairs, hull, cap=4 * len(pairs) - 1)[0] == "CAPACITY"
    bad = pairs.copy()
    bad["b"][3] = len(bodies)
    assert _noalloc(bodies, bad, hull)[0] == "INVALID"
    assert _noalloc(bodies, pairs, hull[:100])[0] == "INVALID"  # a hull range past the pool
    frac = bodies.copy()
    frac["corners"][np.flatnonzero(frac["tag"] == R.HULL)[0], 0, 1] = 41.5
    assert _noalloc(frac, pairs, hull)[0] == "INVALID"
    many = np.zeros(MAX_BODIES + 1, R.GPU_BODY)
    many["orientation"][:, 3] = 1.0
    assert _noalloc(many, np.zeros(1, R.GPU_PAIR), hull)[0] == "CAPACITY"
    assert _noalloc(bodies, np.zeros(MAX_PAIRS + 1, R.GPU_PAIR), hull)[0] == "CAPACITY"
    assert _noalloc(bodies, pairs, np.zeros((MAX_HULL + 1, 3), np.float32))[0] == "CAPACITY"
    # an empty pair list is a no-op
    assert _noalloc(bodies, np.zeros(0, R.GPU_PAIR), hull)[0] == "OK"
    # and the entry point still works after refusing
    st, got = _noalloc(bodies, pairs, hull)
    assert st == "OK" and got.tobytes() == _reference(bodies, pairs, hull).tobytes()

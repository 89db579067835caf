"""Full-size, full-length GPU parity: each benchmark workload at its real size
and step count, the device step against the FP32 oracle bit for bit.

- the 8192-box cube field, 300 steps from the drop (poses, velocities,
  contacts, multipliers and penalties every 50 steps, the pair list at the
  end), with its distance from the FP64 oracle after 300 steps stated;
- the mixed convex field (20000 hulls, tetrahedra and boxes, 128 ropes of 64
  links, 512 springs), 600 steps at dt = 1/120;
- the 4096-scene table stepped as one world, 300 steps, against the oracle
  stepping 256 of its scenes alone (every 16th scene by size, so the sample
  spans 8 to 512 bodies);
- 64 wall throws batched in one world, 120 steps, against the oracle stepping
  each wall alone.

The step protocol is the reference demo's (demos/avbd_box/src/main.rs:189-272,
contact lag 1).
"""

from __future__ import annotations

import time

import numpy as np
import pytest

from diomedes_amd import PhysicsContext, scenes

from .helpers import assert_bits

pytestmark = pytest.mark.gpu

KEYS = ("positions", "orientations", "velocities", "angular_velocities", "prev_velocities")


def _print(*a):
    print(*a, flush=True)


class _SceneContacts:
    """A batched world's contact list split by scene, each scene's contacts
    in list order with scene-local body indices (the container stays
    CONTAINER): the order the oracle gives stepping that scene alone."""

    def __init__(self, gc, offsets):
        self.offsets = np.asarray(offsets, np.int64)
        scene = np.searchsorted(self.offsets, gc["a"].astype(np.int64), side="right") - 1
        order = np.argsort(scene, kind="stable")
        self.gc = gc[order]
        self.bounds = np.searchsorted(scene[order], np.arange(len(self.offsets)))

    def __call__(self, s):
        c = self.gc[self.bounds[s]:self.bounds[s + 1]].copy()
        lo = np.uint32(self.offsets[s])
        c["a"] -= lo
        notc = c["b"] != 0xFFFFFFFF
        c["b"][notc] -= lo
        return c


def _compare_contacts(c, rc, what):
    assert len(c) == len(rc["a"]), f"{what}: {len(c)} contacts vs {len(rc['a'])}"
    assert np.array_equal(c["a"], rc["a"]) and np.array_equal(c["b"], rc["b"]), f"{what}: contact bodies"
    assert_bits(c["lambda_"][:, :3], rc["lambda"], f"{what} lambda")
    assert_bits(c["penalty"][:, :3], rc["penalty"], f"{what} penalty")
    assert_bits(c["r_a"][:, :3], rc["r_a"], f"{what} r_a")
    assert_bits(c["normal"][:, :3], rc["normal"], f"{what} normal")


def _compare_world(ctx, w, what):
    got, ref = ctx.download(), w.state()
    for key in KEYS:
        assert_bits(got[key], ref[key], f"{what} {key}")
    _compare_contacts(ctx.contacts(), w.contacts(), what)
    return got


def test_cube_field_8192_300_steps_bitexact(oracle):
    """The plain cube field (32 x 8 x 32 boxes, I = 4, friction 0.6) for the
    full 300 steps of its benchmark, from the drop to the settled pile."""
    scene = scenes.cube_field(0)
    w = oracle.OracleWorld(scene, 32)
    t0 = time.perf_counter()
    with PhysicsContext(scene.n) as ctx:
        ctx.upload(scene)
        for k in range(6):
            ctx.step(scene.options, 50)
            w.step(nsteps=50)
            got = _compare_world(ctx, w, f"step {50 * (k + 1)}")
            s = ctx.stats()
            _print("cube field 8192 step", 50 * (k + 1), "pairs", s.pairs, "contacts", s.contacts,
                   f"{time.perf_counter() - t0:.1f} s")
        assert s.contacts > 20000, "the field has not settled into a pile"
        pairs = ctx.pairs()
        assert np.array_equal(np.stack([pairs["a"], pairs["b"]], 1).reshape(-1, 2), w.pairs().reshape(-1, 2))
    # distance from the FP64 oracle after 300 steps (stated, not a bound: the
    # contact manifolds are discrete choices, so the trajectories part once
    # boxes collide)
    w64 = oracle.OracleWorld(scene, 64)
    w64.step(nsteps=300)
    p64 = w64.state()["positions"]
    dev = np.linalg.norm(got["positions"] - p64, axis=1)
    scale = np.linalg.norm(p64, axis=1).max()
    _print(f"cube field 8192, 300 steps: |x_fp32 - x_fp64| max {dev.max():.3e} m, median {np.median(dev):.3e} m, "
           f"max relative to the field's extent {dev.max() / scale:.3e}; bodies within 1e-4 relative: "
           f"{np.mean(dev / scale < 1e-4):.3f}")


def test_mixed_field_20000_600_steps_bitexact(oracle):
    """The mixed convex field with ropes and springs for its 600 steps."""
    scene = scenes.mixed_field(0)
    w = oracle.OracleWorld(scene, 32)
    t0 = time.perf_counter()
    with PhysicsContext(scene.n, max_hull_vertices=len(scene.hull), max_constraints=len(scene.constraints)) as ctx:
        ctx.upload(scene)
        for k in range(6):
            ctx.step(scene.options, 100)
            w.step(nsteps=100)
            _compare_world(ctx, w, f"step {100 * (k + 1)}")
            s = ctx.stats()
            _print("mixed 20000 step", 100 * (k + 1), "pairs", s.pairs, "contacts", s.contacts,
                   f"{time.perf_counter() - t0:.1f} s")
        assert s.contacts > 10000


def test_batched_4096_scenes_match_oracle_per_scene(oracle):
    """The full 4096-scene table (507,502 bodies) stepped as one world; 256
    scenes spanning the size range are stepped by the oracle alone and compared
    every 60 steps over the 300 steps of the workload."""
    batch = scenes.batched_scenes(0)
    opts = batch.options
    sizes = np.array([sc.n for sc in batch.scenes])
    order = np.argsort(sizes, kind="stable")
    sample = order[np.arange(0, len(order), len(order) // 256)][:256]
    assert len(sample) == 256 and sizes[sample].min() <= 9 and sizes[sample].max() >= 500
    worlds = {int(s): oracle.OracleWorld(batch.scenes[s], 32) for s in sample}
    t0 = time.perf_counter()
    with PhysicsContext(batch.n, max_pairs=8 * batch.n, max_contacts=16 * batch.n,
                        max_hull_vertices=len(batch.hull)) as ctx:
        ctx.batch_upload(batch)
        for k in range(5):
            ctx.step(opts, 60)
            got = ctx.download()
            gc = ctx.contacts()
            split = _SceneContacts(gc, batch.offsets)
            for s, w in worlds.items():
                w.step(nsteps=60)
                sl = batch.scene_slice(s)
                ref = w.state()
                for key in KEYS:
                    assert_bits(got[key][sl], ref[key], f"scene {s} ({sizes[s]} bodies) step {60 * (k + 1)} {key}")
                _compare_contacts(split(s), w.contacts(), f"scene {s} step {60 * (k + 1)}")
            _print("batched 4096 step", 60 * (k + 1), "contacts", len(gc), f"{time.perf_counter() - t0:.1f} s")


def test_batched_4096_scene_solve_bitexact():
    """The full 4096-scene table over 120 steps with the per-scene solve (one
    launch, bundles of at least 256 bodies) and with the span primal against
    the per-launch schedule with one lane per body, all unfused as the
    autotuner tunes this table: every body bit-identical every 40 steps."""
    batch = scenes.batched_scenes(0)
    opts = batch.options
    res = {}
    runs = {"lanes": dict(primal_lanes=1, primal_packed=0, fuse_dual=0, scene_solve=0),
            "scene": dict(primal_lanes=1, primal_packed=0, fuse_dual=0, scene_solve=256),
            "span": dict(primal_packed=2, fuse_dual=0, scene_solve=0)}
    for name, tuning in runs.items():
        with PhysicsContext(batch.n, max_pairs=8 * batch.n, max_contacts=16 * batch.n,
                            max_hull_vertices=len(batch.hull)) as ctx:
            ctx.set_tuning(**tuning)
            ctx.batch_upload(batch)
            res[name] = []
            for _ in range(3):
                ctx.step(opts, 40)
                res[name].append(ctx.download())
    for name in ("scene", "span"):
        for k in range(3):
            for key in KEYS:
                assert_bits(res[name][k][key], res["lanes"][k][key], f"{name} step {40 * (k + 1)} {key}")
    _print("batched 4096, scene solve and span primal: bit-identical over 120 steps")


def test_wall_throw_64_seeds_120_steps_bitexact(oracle):
    """64 walls of 64 x 32 bricks, each hit by its own projectile, batched in
    one world for the 120 steps of the differentiable workload; each wall
    against the oracle stepping it alone."""
    throws = [scenes.wall_throw(s) for s in range(64)]
    batch = scenes.BatchedScenes(throws, np.zeros((0, 3), np.float32), throws[0].options)
    opts = throws[0].options
    worlds = [oracle.OracleWorld(sc, 32) for sc in throws]
    t0 = time.perf_counter()
    with PhysicsContext(batch.n, max_pairs=8 * batch.n, max_contacts=16 * batch.n) as ctx:
        ctx.batch_upload(batch)
        for k in range(3):
            ctx.step(opts, 40)
            got = ctx.download()
            gc = ctx.contacts()
            split = _SceneContacts(gc, batch.offsets)
            for s, w in enumerate(worlds):
                w.step(nsteps=40)
                sl = batch.scene_slice(s)
                ref = w.state()
                for key in KEYS:
                    assert_bits(got[key][sl], ref[key], f"wall {s} step {40 * (k + 1)} {key}")
                _compare_contacts(split(s), w.contacts(), f"wall {s} step {40 * (k + 1)}")
            _print("wall throws step", 40 * (k + 1), "contacts", len(gc), f"{time.perf_counter() - t0:.1f} s")

"""GPU parity: the HIP kernels against the FP32 oracle, through the C ABI.

Bit-exact is the bar for everything on the FP32 path: pair lists (integer),
AABBs, narrow-phase contacts, AVBD poses / velocities / multipliers, and whole
multi-step trajectories of the fused device step.  Both sides evaluate the
same float operations in the same order with no fused multiply-adds.
"""

from __future__ import annotations

import numpy as np
import pytest

from diomedes_amd import PhysicsContext, scenes
from diomedes_amd import records as R
from diomedes_amd.physics import AvbdOptions

from .helpers import assert_bits, contacts_to_records, primal_tuning

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


# ------------------------------------------------------------ broad phase
@pytest.mark.parametrize("blanes", [1, 2, 4, 8, 16, 32])
@pytest.mark.parametrize("nx,ny,nz,seed", [(4, 2, 4, 1), (8, 2, 8, 2), (32, 8, 32, 0), (1, 1, 3, 5), (16, 4, 16, 6),
                                          (64, 10, 64, 6)])
def test_broad_pairs_match_dense_sweep(oracle, nx, ny, nz, seed, blanes):
    _print("broad", nx, ny, nz, seed, blanes)
    scene = scenes.cube_field(seed, nx=nx, ny=ny, nz=nz, spacing=1.2)
    with PhysicsContext(scene.n, max_pairs=64 * scene.n) as ctx:
        ctx.set_tuning(broad_lanes=blanes)
        ctx.broad_phase_submit(scene.bodies)
        pairs = ctx.broad_phase_pairs()
        aabbs = ctx.broad_phase_aabbs(scene.n)
    ref = oracle.broad_pairs(scene.bodies)
    ref_aabb = oracle.aabbs(scene.bodies)
    assert_bits(aabbs[:, [0, 1, 2, 4, 5, 6]], ref_aabb, "aabbs")
    got = np.stack([pairs["a"], pairs["b"]], axis=1)
    assert got.shape == ref.shape, (got.shape, ref.shape)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("blanes", [8, 16, 32])
def test_broad_degenerate_inputs(oracle, blanes):
    # one body, two coincident bodies, many identical bodies (equal Morton
    # codes; every node overlaps every query, the traversal stack's worst case)
    for n in (1, 2, 50, 600):
        scene = scenes.cube_field(3, nx=n, ny=1, nz=1, spacing=0.0)
        with PhysicsContext(max(n, 1), max_pairs=n * n + 16) as ctx:
            ctx.set_tuning(broad_lanes=blanes)
            ctx.broad_phase_submit(scene.bodies)
            pairs = ctx.broad_phase_pairs()
        ref = oracle.broad_pairs(scene.bodies)
        assert np.array_equal(np.stack([pairs["a"], pairs["b"]], 1).reshape(-1, 2), ref.reshape(-1, 2))
        assert len(ref) == n * (n - 1) // 2


def test_broad_capacity_reports_required_size(oracle):
    from diomedes_amd._lib import DioError

    scene = scenes.cube_field(3, nx=20, ny=1, nz=1, spacing=0.0)  # 190 pairs
    with PhysicsContext(scene.n, max_pairs=64) as ctx:
        ctx.broad_phase_submit(scene.bodies)
        with pytest.raises(DioError) as e:
            ctx.broad_phase_count()
        assert e.value.status == "CAPACITY" and e.value.required == 190


# ------------------------------------------------------------ narrow phase
def _settled_field(oracle, nx=6, ny=3, nz=6, steps=40, seed=4):
    scene = scenes.cube_field(seed, nx=nx, ny=ny, nz=nz)
    w = oracle.OracleWorld(scene, 32)
    w.step(nsteps=steps)
    st = w.state()
    bodies = scene.bodies.copy()
    bodies["position"][:, :3] = st["positions"].astype(np.float32)
    bodies["orientation"] = st["orientations"].astype(np.float32)
    return scene, w, bodies


def test_narrow_box_manifolds_bitexact(oracle):
    scene, _, bodies = _settled_field(oracle)
    pairs = oracle.broad_pairs(bodies)
    _print("narrow pairs", len(pairs))
    assert len(pairs) > 20
    p = np.zeros(len(pairs), R.GPU_PAIR)
    p["a"], p["b"] = pairs[:, 0], pairs[:, 1]
    with PhysicsContext(scene.n) as ctx:
        raw = ctx.narrow_phase_run(bodies, p)
    assert len(raw) == 4 * len(pairs)
    total = 0
    for i, (a, b) in enumerate(pairs):
        ref = oracle.narrow_pair(bodies[a], bodies[b])
        slots = raw[4 * i:4 * i + 4]
        assert int(np.sum(slots["valid"])) == len(ref), (i, a, b)
        for k, c in enumerate(ref):
            s = slots[k]
            assert s["valid"] == 1 and s["a"] == a and s["b"] == b
            assert_bits(s["normal"][:3], c["normal"], f"normal {i}/{k}")
            assert_bits(s["depth"], c["depth"], f"depth {i}/{k}")
            assert_bits(s["point_a"][:3], c["point_a"], f"pa {i}/{k}")
            assert_bits(s["point_b"][:3], c["point_b"], f"pb {i}/{k}")
            total += 1
        assert np.all(slots["valid"][len(ref):] == 0)
    assert total > 0


# ------------------------------------------------------------ AVBD solve
@pytest.mark.parametrize("lanes,fuse", [(1, 0), (4, 1), (8, 0), (16, 1), (32, 0), (32, 1), (1, 1), (64, 0), (64, 1)])
def test_avbd_granular_solve_bitexact(oracle, lanes, fuse):
    scene, w, _ = _settled_field(oracle, steps=30)
    # the oracle world's current contacts / CSR / state feed one granular solve
    c = w.contacts()
    off, idx = w.csr()
    st = w.state()
    state = scene.state.copy()
    state["vel"][:, :3] = st["velocities"]
    state["ang"][:, :3] = st["angular_velocities"]
    state["prev_vel"][:, :3] = st["prev_velocities"]
    pos = st["positions"].astype(np.float32)
    rot = st["orientations"].astype(np.float32)
    con = contacts_to_records(c)
    opts = scene.options
    _print("granular avbd contacts", len(con))
    ref = oracle.avbd_run(state, pos, rot, con, scene.constraints, off, idx, opts.run_options())
    with PhysicsContext(scene.n) as ctx:
        ctx.set_tuning(primal_lanes=lanes, fuse_dual=fuse)
        ctx.gpu_physics_submit(state, pos, rot, con, scene.constraints, off, idx, opts)
        got = ctx.gpu_physics_read()
    for k in ("positions", "orientations", "velocities", "angular_velocities", "prev_velocities", "lambda", "penalty"):
        assert_bits(got[k], ref[k], k)


@pytest.mark.parametrize("spokes", [300, 700])
@pytest.mark.parametrize("lanes", ["packed", 32])
def test_hub_with_hundreds_of_items_bitexact(oracle, spokes, lanes):
    """A hub body held by springs to `spokes` satellites: its workgroup's
    joined item list passes 256 entries, so the packed primal builds it in
    several rounds and, in iterations >= 1, takes the entries past its item
    table from the CSR; the grouped primal takes the hub in many chunks."""
    from diomedes_amd.physics import AvbdBody, Scene, Spring

    rng = np.random.default_rng(spokes)
    bodies = [AvbdBody.cube((0.0, 4.0, 0.0), 0.5, 20.0)]
    cons = []
    for k in range(spokes):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        bodies.append(AvbdBody.cube(tuple(np.array([0.0, 4.0, 0.0]) + 2.5 * d), 0.08, 0.05))
        cons.append(Spring(0, k + 1, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 40.0, 2.4))
    opts = AvbdOptions(gravity=(0.0, -9.81, 0.0), iterations=6)
    scene = Scene.from_bodies(bodies, cons, scenes.floor_container(20.0), opts)
    w = oracle.OracleWorld(scene, 32)
    with PhysicsContext(scene.n, max_constraints=len(cons)) as ctx:
        ctx.set_tuning(**primal_tuning(lanes))
        ctx.upload(scene)
        for k in range(3):
            w.step(opts, nsteps=15)
            ctx.step(opts, 15)
            got, ref = ctx.download(), w.state()
            for key in ("positions", "orientations", "velocities", "angular_velocities"):
                assert_bits(got[key], ref[key], f"{lanes} {spokes} spokes, step {15 * (k + 1)} {key}")
        _print("hub", spokes, lanes, "contacts", ctx.stats().contacts)


@pytest.mark.parametrize("lanes", [1, 8, 64, "packed", "span"])
def test_avbd_granular_joints_bitexact(oracle, lanes):
    """Springs, rods, a ball socket and a rope (the constraints demo's set,
    demos/constraints/src/main.rs:110-146) through one solve."""
    from diomedes_amd.physics import AvbdBody, BallSocket, Rod, Rope, Scene, Spring

    bodies = [
        AvbdBody.cube((-7.0, 1.6, 0.0), 0.4, 0.0), AvbdBody.cube((-7.0, 0.0, 0.0), 0.3, 1.0),
        AvbdBody.cube((-2.5, 1.6, 0.0), 0.4, 0.0), AvbdBody.cube((-1.7, 0.2, 0.0), 0.3, 1.0),
        AvbdBody.cube((-2.2, -1.0, 0.0), 0.3, 1.0), AvbdBody.cube((2.2, 1.7, 0.0), 0.4, 0.0),
        AvbdBody.cube((2.2, 2.05, 0.0), 0.3, 1.0), AvbdBody.cube((6.5, 1.7, 0.0), 0.4, 0.0),
        AvbdBody.cube((6.5, -0.2, 0.0), 0.3, 1.0),
    ]
    for i, v in ((1, (9.2, 4.4, 5.0)), (3, (-5.4, 4.0, 5.0)), (4, (5.7, 4.2, 4.0)), (6, (5.0, 4.0, 4.8)),
                 (8, (-5.8, 5.5, 4.0))):
        bodies[i].velocity = v
    cons = [Spring(0, 1, (0, 0, 0), (0, 0, 0), 12.0, 1.6), Rod(2, 3, (0, 0, 0), (0, 0, 0), 1.6),
            Rod(3, 4, (0, 0, 0), (0, 0, 0), 1.4), BallSocket(5, 6, (0, 0, 0), (0.0, -0.3, 0.0)),
            Rope(7, 8, (0, 0, 0), (0, 0, 0), 1.5)]
    opts = AvbdOptions(gravity=(0.0, -3.5, 0.0), iterations=16)
    scene = Scene.from_bodies(bodies, cons, None, opts)
    w = oracle.OracleWorld(scene, 32)
    with PhysicsContext(scene.n) as ctx:
        ctx.set_tuning(**primal_tuning(lanes))
        ctx.upload(scene)
        for k in range(4):
            w.step(opts, nsteps=30)
            ctx.step(opts, 30)
            got, ref = ctx.download(), w.state()
            for key in ("positions", "orientations", "velocities", "angular_velocities", "prev_velocities"):
                assert_bits(got[key], ref[key], f"step {30 * (k + 1)} {key}")


# ------------------------------------------------------------ fused step
def _compare_world(ctx, w, step, check_contacts=True):
    got, ref = ctx.download(), w.state()
    for key in ("positions", "orientations", "velocities", "angular_velocities", "prev_velocities"):
        assert_bits(got[key], ref[key], f"step {step} {key}")
    if check_contacts:
        gc = ctx.contacts()
        rc = w.contacts()
        assert len(gc) == len(rc["a"]), f"step {step}: {len(gc)} contacts vs {len(rc['a'])}"
        assert np.array_equal(gc["a"], rc["a"]) and np.array_equal(gc["b"], rc["b"])
        assert_bits(gc["lambda_"][:, :3], rc["lambda"], f"step {step} lambda")
        assert_bits(gc["penalty"][:, :3], rc["penalty"], f"step {step} penalty")
        assert_bits(gc["r_a"][:, :3], rc["r_a"], f"step {step} r_a")


@pytest.mark.parametrize("limit", ["pairs", "contacts"])
def test_world_step_capacity(oracle, limit):
    """A fused step that needs more candidate pairs than max_pairs, or more
    contacts than max_contacts: every step before it stays bit-exact, that
    step returns CAPACITY with the count the oracle finds at that step and
    names the limit, later calls report it again, and an upload clears it.
    Every device write stays within the limits (the lists are truncated)."""
    from diomedes_amd._lib import DioError

    scene = scenes.cube_field(4, nx=6, ny=3, nz=6)  # 45 pairs at first, growing as the field collapses
    cap = 50 if limit == "pairs" else 40
    kw = {"max_pairs": cap} if limit == "pairs" else {"max_contacts": cap}
    w = oracle.OracleWorld(scene, 32)
    with PhysicsContext(scene.n, **kw) as ctx:
        ctx.upload(scene)
        for step in range(1, 81):
            w.step(nsteps=1)
            need = len(w.pairs()) if limit == "pairs" else len(w.contacts()["a"])
            if need > cap:
                break
            ctx.step(scene.options, 1)
            _compare_world(ctx, w, step)
        assert need > cap, "the scene never reached the limit"
        _print(limit, "limit", cap, "first exceeded at step", step, "needing", need)
        with pytest.raises(DioError) as e:
            ctx.step(scene.options, 1)
        assert e.value.status == "CAPACITY" and e.value.required == need
        assert f"max_{limit} ({cap})" in str(e.value)
        with pytest.raises(DioError) as e:
            ctx.step(scene.options, 2)
        assert e.value.status == "CAPACITY"
        ctx.upload(scene)
        ctx.step(scene.options, 1)


def test_sphere_heap_bitexact(oracle):
    """A world of spheres only (no pair reaches GJK, so the step leaves out
    the EPA launches): 400 spheres of two radii dropped into a container."""
    from diomedes_amd.physics import AvbdBody, AvbdOptions, Scene

    rng = np.random.default_rng(11)
    bodies = []
    for i in range(400):
        x, z = (i % 20) * 1.1 - 10.0, (i // 20) * 1.1 - 10.0
        r = 0.3 if i % 3 else 0.5
        bodies.append(AvbdBody.sphere((x + rng.uniform(-0.05, 0.05), 1.0 + (i % 7) * 1.2, z), r, 1.0))
        bodies[-1].velocity = tuple(rng.normal(0.0, 1.0, 3))
    scene = Scene.from_bodies(bodies, (), scenes.floor_container(14.0), AvbdOptions(iterations=6))
    w = oracle.OracleWorld(scene, 32)
    with PhysicsContext(scene.n) as ctx:
        ctx.upload(scene)
        for k in range(4):
            ctx.step(scene.options, 15)
            w.step(nsteps=15)
            _compare_world(ctx, w, 15 * (k + 1))
            _print("sphere heap step", 15 * (k + 1), "pairs", ctx.stats().pairs, "contacts", ctx.stats().contacts)
        assert ctx.stats().contacts > 200


def test_box_stack_300_steps_bitexact(oracle):
    scene = scenes.box_stack(0)
    w = oracle.OracleWorld(scene, 32)
    with PhysicsContext(scene.n) as ctx:
        ctx.upload(scene)
        for k in range(10):
            ctx.step(scene.options, 30)
            w.step(nsteps=30)
            _compare_world(ctx, w, 30 * (k + 1))
        pairs = ctx.pairs()
        ref_pairs = w.pairs()
        assert np.array_equal(np.stack([pairs["a"], pairs["b"]], 1).reshape(-1, 2), ref_pairs.reshape(-1, 2))


@pytest.mark.parametrize(
    "lag,lanes,blanes,resort,fuse",
    [(1, 8, 4, 4, 0), (0, 8, 4, 4, 1), (1, 4, 2, 1, 0), (1, 16, 8, 1000, 1), (1, 1, 1, 4, 1), (0, 4, 8, 1000, 0),
     (1, 32, 1, 4, 1), (1, 32, 8, 4, 0), (1, 64, 8, 4, 1), (1, 32, 16, 16, 1), (0, 32, 32, 4, 1),
     (1, "packed", 8, 4, 1), (1, "packed", 16, 16, 0), (0, "packed", 4, 1000, 1), (0, "packed", 1, 4, 0),
     (1, "span", 8, 4, 1), (0, "span", 4, 1000, 0), (1, "span", 16, 16, 0)],
)
def test_cube_field_small_bitexact(oracle, lag, lanes, blanes, resort, fuse):
    """resort: Morton re-sort interval of the fused step; 1000 refits the first
    step's order for the whole run, which must not change the pair set.
    fuse: dual updates applied inside the next primal launch."""
    scene = scenes.cube_field(7, nx=8, ny=2, nz=8)
    w = oracle.OracleWorld(scene, 32)
    with PhysicsContext(scene.n) as ctx:
        ctx.set_tuning(**primal_tuning(lanes), broad_lanes=blanes, resort_interval=resort, fuse_dual=fuse)
        ctx.upload(scene)
        for k in range(6):
            ctx.step(scene.options, 20, lag=lag)
            w.step(nsteps=20, lag=lag)
            _compare_world(ctx, w, 20 * (k + 1))


def test_contact_lag_switches_bitexact(oracle):
    """Lag 1, then lag 0, then lag 1 again on one world: after a lag-0 step the
    next lag-1 step builds its contacts from the narrow output that step found
    on the poses at its start, as if it had run with lag 1 (the oracle does
    the same), and graph capture follows each switch."""
    scene = scenes.cube_field(9, nx=8, ny=2, nz=8)
    w = oracle.OracleWorld(scene, 32)
    done = 0
    with PhysicsContext(scene.n) as ctx:
        ctx.upload(scene)
        for lag, n in ((1, 25), (0, 7), (1, 1), (0, 1), (1, 20), (0, 10), (1, 15)):
            ctx.step(scene.options, n, lag=lag)
            w.step(nsteps=n, lag=lag)
            done += n
            _compare_world(ctx, w, done)
            pairs = ctx.pairs()
            assert np.array_equal(np.stack([pairs["a"], pairs["b"]], 1).reshape(-1, 2), w.pairs().reshape(-1, 2))


def test_refit_carried_by_the_contact_build_at_depth_six(oracle):
    """More than 32768 bodies (an 8-ary tree of depth 6, whose top levels take
    a launch of their own after the chunk refit): with lag 1 on one stream the
    refit steps carry the chunk refit on the contact build's first launch;
    every step equals the oracle's, and the pairs equal those of the same run
    with the broad phase on the second stream (a separate refit launch)."""
    scene = scenes.cube_field(12, nx=64, ny=10, nz=64, spacing=1.6)
    assert scene.n > 32768
    w = oracle.OracleWorld(scene, 32)
    runs = {}
    for overlap in (0, 1):
        with PhysicsContext(scene.n) as ctx:
            ctx.set_tuning(resort_interval=4, overlap_broad=overlap)
            ctx.upload(scene)
            got = []
            for k in range(3):
                ctx.step(scene.options, 3)
                if overlap == 0:
                    w.step(nsteps=3)
                    _compare_world(ctx, w, 3 * (k + 1))
                got.append((ctx.download()["positions"].copy(), ctx.pairs().copy()))
            runs[overlap] = got
            _print("depth-6 field, overlap", overlap, "pairs", ctx.stats().pairs, "contacts", ctx.stats().contacts)
    for (p0, q0), (p1, q1) in zip(runs[0], runs[1]):
        assert_bits(p1, p0, "positions with and without the carried refit")
        assert np.array_equal(q0, q1)


@pytest.mark.parametrize("order,lanes,fuse", [(1, 32, 1), (1, 1, 0), (1, 8, 1), (0, 32, 1), (0, 1, 0),
                                              (1, "packed", 1), (0, "packed", 1), (1, "packed", 0),
                                              (1, "span", 1), (0, "span", 0)])
def test_primal_body_order_bitexact(oracle, order, lanes, fuse):
    """primal_order 1 (the default): the Jacobi primal takes the bodies in the
    broad phase's Morton order, one contiguous range per XCD; 0: in index
    order.  Either way every body's result is the oracle's."""
    scene = scenes.cube_field(13, nx=12, ny=3, nz=12)
    w = oracle.OracleWorld(scene, 32)
    with PhysicsContext(scene.n) as ctx:
        assert ctx.tuning()["primal_order"] == 1
        ctx.set_tuning(**primal_tuning(lanes), fuse_dual=fuse, primal_order=order)
        assert ctx.tuning()["primal_order"] == order
        ctx.upload(scene)
        for k in range(4):
            ctx.step(scene.options, 20)
            w.step(nsteps=20)
            _compare_world(ctx, w, 20 * (k + 1))


def test_batched_scenes_in_morton_order_bitexact(oracle):
    """The Morton-ordered primal over a scene table: each scene as the oracle
    steps it alone."""
    batch = scenes.batched_scenes(4, num_scenes=12, n_min=8, n_max=200)
    worlds = [oracle.OracleWorld(sc, 32) for sc in batch.scenes]
    with PhysicsContext(batch.n, max_hull_vertices=len(batch.hull)) as ctx:
        ctx.set_tuning(primal_lanes=1, fuse_dual=0, primal_order=1)
        ctx.batch_upload(batch)
        for k in range(2):
            ctx.step(batch.options, 25)
            got = ctx.download()
            for s, w in enumerate(worlds):
                w.step(nsteps=25)
                sl = batch.scene_slice(s)
                for key in ("positions", "orientations", "velocities", "angular_velocities"):
                    assert_bits(got[key][sl], w.state()[key], f"scene {s} step {25 * (k + 1)} {key}")


def test_autotune_on_a_batched_table_keeps_parity(oracle):
    """The autotuner on a scene table: its broad-phase timing keeps to each
    scene's bodies (scenes overlap in space), it leaves the world unchanged,
    and the tuned steps stay each scene's own."""
    import time

    batch = scenes.batched_scenes(6, num_scenes=40, n_min=8, n_max=120)
    worlds = [oracle.OracleWorld(sc, 32) for sc in batch.scenes]
    with PhysicsContext(batch.n, max_hull_vertices=len(batch.hull)) as ctx:
        ctx.batch_upload(batch)
        ctx.step(batch.options, 15)
        t0 = time.perf_counter()
        res = ctx.autotune(batch.options, reps=5)
        _print("batched autotune", round(time.perf_counter() - t0, 3), "s:", res["primal_lanes"], res["broad_lanes"],
               res["fuse_dual"], res["primal_order"])
        ctx.step(batch.options, 15)
        got = ctx.download()
        for s, w in enumerate(worlds):
            w.step(nsteps=30)
            sl = batch.scene_slice(s)
            for key in ("positions", "orientations", "velocities", "angular_velocities"):
                assert_bits(got[key][sl], w.state()[key], f"scene {s} step 30 {key}")


def test_autotune_keeps_state_and_parity(oracle):
    """The autotuner times the primal shapes without changing the world."""
    scene = scenes.cube_field(11, nx=8, ny=2, nz=8)
    w = oracle.OracleWorld(scene, 32)
    with PhysicsContext(scene.n) as ctx:
        ctx.upload(scene)
        ctx.step(scene.options, 40)
        w.step(nsteps=40)
        res = ctx.autotune(scene.options, reps=5)
        _print("autotune", res)
        assert res["primal_lanes"] in (1, 4, 8, 16, 32, 64)
        assert res["broad_lanes"] in (1, 2, 4, 8, 16, 32)
        assert res["fuse_dual"] in (0, 1)
        assert res["primal_order"] in (0, 1) and ctx.tuning()["primal_order"] == res["primal_order"]
        assert res["colour_lanes"] in (1, 4, 8, 16, 32, 64) and ctx.tuning()["colour_lanes"] == res["colour_lanes"]
        assert all(v > 0 for v in res["colour_us_per_iteration"].values())
        _compare_world(ctx, w, 40)
        # coloured steps at the chosen colour lanes keep parity too
        w.set_mode(1)
        ctx.step(scene.options, 12, mode=1)
        w.step(nsteps=12)
        _compare_world(ctx, w, 52)
        from diomedes_amd._lib import DioError

        with pytest.raises(DioError):
            ctx.set_tuning(colour_lanes=3)
        w.set_mode(0)
        ctx.step(scene.options, 20)
        w.step(nsteps=20)
        _compare_world(ctx, w, 72)


def test_cube_field_8192_first_steps_bitexact(oracle):
    scene = scenes.cube_field(0)
    w = oracle.OracleWorld(scene, 32)
    with PhysicsContext(scene.n) as ctx:
        ctx.upload(scene)
        for k in range(3):
            ctx.step(scene.options, 1)
            w.step(nsteps=1)
            _compare_world(ctx, w, k + 1)
            s = ctx.stats()
            _print("8192 step", k, "pairs", s.pairs, "contacts", s.contacts)


def test_mixed_field_20000_first_steps_bitexact(oracle):
    """The full mixed workload: 20000 hulls, tetrahedra and boxes, 128 ropes of
    64 links and 512 springs, GJK + EPA and joints."""
    scene = scenes.mixed_field(0)
    w = oracle.OracleWorld(scene, 32)
    with PhysicsContext(scene.n, max_hull_vertices=len(scene.hull), max_constraints=len(scene.constraints)) as ctx:
        ctx.upload(scene)
        done = 0
        for n in (30, 10, 1, 1):  # the bodies start apart; collisions begin after about 15 steps
            ctx.step(scene.options, n)
            w.step(nsteps=n)
            done += n
            _compare_world(ctx, w, done)
            s = ctx.stats()
            _print("mixed 20000 step", done, "pairs", s.pairs, "contacts", s.contacts)
        assert s.contacts > 1000


# ------------------------------------------------------------ golden fixtures
def test_box_stack_matches_golden_fixture():
    """The committed FP32 trajectory (tests/golden/box_stack_s0.npz), bit-exact."""
    from pathlib import Path

    g = np.load(Path(__file__).resolve().parent / "golden" / "box_stack_s0.npz")
    scene = scenes.box_stack(0)
    every = int(g["every"])
    with PhysicsContext(scene.n) as ctx:
        ctx.upload(scene)
        for s in range(int(g["steps"]) // every):
            ctx.step(scene.options, every)
            got = ctx.download()
            for k in ("positions", "orientations", "velocities", "angular_velocities"):
                assert_bits(got[k], g[f"fp32_{k}"][s], f"step {every * (s + 1)} {k}")
        gc = ctx.contacts()
        assert np.array_equal(np.stack([gc["a"], gc["b"]], 1), g["fp32_contact_ab"])
        assert_bits(gc["lambda_"][:, :3], g["fp32_contact_lambda"], "lambda")
        # error against the FP64 oracle, stated
        p64 = g["fp64_positions"][-1]
        _print("box stack: max |x_gpu - x_fp64| after 300 steps =", float(np.max(np.abs(got["positions"] - p64))))


# ------------------------------------------------------------ convex shapes
def _convex_pair_bodies(seed: int, count: int):
    """Random overlapping pairs of tetrahedra, hulls, spheres and boxes."""
    from diomedes_amd.physics import AvbdBody, Shape, gpu_bodies

    rng = np.random.default_rng(seed)
    ico = scenes.icosphere(1).astype(np.float32)
    hull = []
    bodies = []

    def shape(kind):
        if kind == 0:
            return Shape.tetrahedron(scenes.regular_tetrahedron(rng.uniform(0.5, 1.0)) @ _rot(rng).T)
        if kind == 1:
            r = rng.uniform(0.3, 0.6)
            hull.append(ico * np.float32(r))
            return Shape.hull(42 * (len(hull) - 1), 42)
        if kind == 2:
            return Shape.sphere(rng.uniform(0.2, 0.6))
        return Shape.box(rng.uniform(0.2, 0.5, 3))

    for k in range(count):
        ka, kb = rng.integers(0, 4, 2)
        if ka == 2 and kb == 2 or {ka, kb} == {3}:
            ka = 0
        q = rng.normal(size=(2, 4))
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        pa = rng.normal(size=3)
        bodies.append(AvbdBody(shape(ka), tuple(pa), tuple(q[0])))
        bodies.append(AvbdBody(shape(kb), tuple(pa + rng.normal(size=3) * 0.4), tuple(q[1])))
    hull = np.concatenate(hull) if hull else np.zeros((0, 3), np.float32)
    return gpu_bodies(bodies), hull


def _rot(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q


@pytest.mark.parametrize("seed", [0, 1])
def test_narrow_convex_gjk_epa_bitexact(oracle, seed):
    bodies, hull = _convex_pair_bodies(seed, 300)
    n = len(bodies)
    pairs = np.zeros(n // 2, R.GPU_PAIR)
    pairs["a"] = np.arange(0, n, 2)
    pairs["b"] = np.arange(1, n, 2)
    with PhysicsContext(n, max_hull_vertices=max(len(hull), 1)) as ctx:
        if len(hull):
            from diomedes_amd.physics import Scene
            from diomedes_amd import records as RR

            st = np.zeros(n, RR.BODY_STATE)
            ctx.upload(Scene(bodies, st, hull=hull))  # the hull pool is shared with the granular pipeline
        raw = ctx.narrow_phase_run(bodies, pairs)
    hits = 0
    for i in range(len(pairs)):
        a, b = int(pairs[i]["a"]), int(pairs[i]["b"])
        ref = oracle.narrow_pair(bodies[a], bodies[b], hull=hull)
        slots = raw[4 * i:4 * i + 4]
        assert int(np.sum(slots["valid"])) == len(ref), (i, bodies[a]["tag"], bodies[b]["tag"])
        for k, c in enumerate(ref):
            assert_bits(slots[k]["normal"][:3], c["normal"], f"normal {i}")
            assert_bits(slots[k]["depth"], c["depth"], f"depth {i}")
            assert_bits(slots[k]["point_a"][:3], c["point_a"], f"pa {i}")
            assert_bits(slots[k]["point_b"][:3], c["point_b"], f"pb {i}")
            hits += 1
    _print("convex contacts", hits)
    assert hits > 50


@pytest.mark.parametrize("level", [3, 2])
def test_narrow_epa_large_polytopes_bitexact(oracle, level):
    """Deeply overlapping icosphere hulls of 642 (level 3) or 162 (level 2)
    vertices: EPA polytopes outgrow the first pass (64 vertices, 128 faces)
    and many reach the 256-vertex cap of epa.rs; the retried pairs must give
    the full-size result.  The level-2 pool (324 vertices) is staged in LDS,
    the level-3 pool (1284) is read in place; both have hulls of more than 64
    vertices, whose wave-wide support takes the least index."""
    from diomedes_amd.physics import AvbdBody, Shape, gpu_bodies
    from diomedes_amd import records as RR
    from diomedes_amd.physics import Scene

    rng = np.random.default_rng(7)
    ico = scenes.icosphere(level).astype(np.float32)
    nv = len(ico)
    hull = np.concatenate([ico * np.float32(0.5), ico * np.float32(0.7)])
    bodies = []
    for k in range(64):
        q = rng.normal(size=(2, 4))
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        pa = rng.normal(size=3) * 3.0
        off = rng.normal(size=3)
        off *= rng.uniform(0.05, 0.9) / np.linalg.norm(off)
        bodies.append(AvbdBody(Shape.hull(0, nv), tuple(pa), tuple(q[0])))
        bodies.append(AvbdBody(Shape.hull(nv, nv), tuple(pa + off), tuple(q[1])))
    bodies = gpu_bodies(bodies)
    n = len(bodies)
    pairs = np.zeros(n // 2, RR.GPU_PAIR)
    pairs["a"] = np.arange(0, n, 2)
    pairs["b"] = np.arange(1, n, 2)
    with PhysicsContext(n, max_hull_vertices=len(hull)) as ctx:
        ctx.upload(Scene(bodies, np.zeros(n, RR.BODY_STATE), hull=hull))
        raw = ctx.narrow_phase_run(bodies, pairs)
    hits = 0
    for i in range(len(pairs)):
        ref = oracle.narrow_pair(bodies[2 * i], bodies[2 * i + 1], hull=hull)
        slots = raw[4 * i:4 * i + 4]
        assert int(np.sum(slots["valid"])) == len(ref), i
        for k, c in enumerate(ref):
            assert_bits(slots[k]["normal"][:3], c["normal"], f"normal {i}")
            assert_bits(slots[k]["depth"], c["depth"], f"depth {i}")
            assert_bits(slots[k]["point_a"][:3], c["point_a"], f"pa {i}")
            assert_bits(slots[k]["point_b"][:3], c["point_b"], f"pb {i}")
            hits += 1
    _print("large-polytope EPA contacts", hits)
    assert hits >= len(pairs) // 2  # the rest end EPA's 64 iterations without converging: no contact


def test_mixed_field_small_bitexact(oracle):
    """Hulls, tetrahedra, boxes, ropes and springs through the fused step."""
    scene = scenes.mixed_field(5, n=300, ropes=4, rope_links=8, springs=16, extent=8.0)
    w = oracle.OracleWorld(scene, 32)
    with PhysicsContext(scene.n, max_hull_vertices=len(scene.hull), max_constraints=len(scene.constraints)) as ctx:
        ctx.upload(scene)
        for k in range(4):
            ctx.step(scene.options, 15)
            w.step(nsteps=15)
            _compare_world(ctx, w, 15 * (k + 1))
            _print("mixed step", 15 * (k + 1), "contacts", ctx.stats().contacts)


@pytest.mark.parametrize("resort", [1, 8])
def test_wall_throw_small_bitexact(oracle, resort):
    """A fast projectile through a brick wall: the tree is refit over a stale
    Morton order between re-sorts while bodies scatter."""
    scene = scenes.wall_throw(2, columns=16, rows=8)
    w = oracle.OracleWorld(scene, 32)
    with PhysicsContext(scene.n) as ctx:
        ctx.set_tuning(resort_interval=resort)
        ctx.upload(scene)
        for k in range(4):
            ctx.step(scene.options, 15)
            w.step(nsteps=15)
            _compare_world(ctx, w, 15 * (k + 1))
            _print("wall step", 15 * (k + 1), "pairs", ctx.stats().pairs, "contacts", ctx.stats().contacts)


def test_options_change_between_calls(oracle):
    """The final launch of a step prepares the next step's init for the same
    dt and gravity; a call with other options, an autotune, or a coloured step
    in between must not reuse it."""
    import dataclasses

    scene = scenes.cube_field(3, nx=5, ny=2, nz=5)
    w = oracle.OracleWorld(scene, 32)
    o1 = scene.options
    o2 = dataclasses.replace(o1, dt=1.0 / 90.0)
    o3 = dataclasses.replace(o1, gravity=(0.5, -9.0, 0.0))
    with PhysicsContext(scene.n) as ctx:
        ctx.upload(scene)
        for k, o in enumerate((o1, o2, o2, o3, o1, o1)):
            ctx.step(o, 7)
            w.step(o, nsteps=7)
            _compare_world(ctx, w, 7 * (k + 1))
            if k == 1:
                ctx.autotune(o, reps=2)
                ctx.set_tuning(primal_lanes=32, broad_lanes=8, fuse_dual=1)
            if k == 3:
                w.set_mode(1)
                ctx.step(o, 3, mode=1)
                w.step(o, nsteps=3)
                w.set_mode(0)


@pytest.mark.parametrize("name", ["field", "mixed"])
def test_overlapped_broad_phase_bitexact(oracle, name):
    """overlap_broad: the next step's broad and narrow phases on a second
    stream beside the contact build; the step stays the oracle's, bit for bit
    (graph-captured, and switched off and on between calls)."""
    if name == "field":
        scene = scenes.cube_field(5, nx=10, ny=3, nz=10)
    else:
        scene = scenes.mixed_field(4, n=600, ropes=4, rope_links=8, springs=16)
    w = oracle.OracleWorld(scene, 32)
    with PhysicsContext(scene.n, max_hull_vertices=len(scene.hull), max_constraints=max(16, len(scene.constraints))) as ctx:
        ctx.set_tuning(overlap_broad=1)
        assert ctx.tuning()["overlap_broad"] == 1
        ctx.upload(scene)
        for k, overlap in enumerate((1, 1, 0, 1)):
            ctx.set_tuning(overlap_broad=overlap)
            ctx.step(scene.options, 12)
            w.step(nsteps=12)
            _compare_world(ctx, w, 12 * (k + 1))
            _print(name, "overlap", overlap, "step", 12 * (k + 1), "pairs", ctx.stats().pairs, "contacts",
                   ctx.stats().contacts)


@pytest.mark.parametrize("fuse,lanes", [(0, 32), (1, 32), (1, 1)])
def test_static_contacts_bitexact(oracle, fuse, lanes):
    """Static bodies in contact with each other and with dynamic bodies: in the
    fused schedule a static body runs no primal, so its contacts' dual updates
    are stored by the dynamic side, and a static-static contact takes all its
    updates in the final launch."""
    from diomedes_amd.physics import AvbdBody, Scene

    bodies = [
        AvbdBody.cube((0.0, 0.5, 0.0), 0.5, 0.0),      # static
        AvbdBody.cube((0.8, 0.5, 0.1), 0.5, 0.0),      # static, overlapping the first
        AvbdBody.cube((0.1, 1.45, 0.0), 0.45, 2.0),    # resting on both
        AvbdBody.cube((1.6, 0.45, 0.0), 0.45, 1.0),    # dynamic b > static a
        AvbdBody.cube((-1.2, 1.5, 0.3), 0.4, 1.0),
    ]
    bodies[4].velocity = (3.0, -1.0, 0.0)
    opts = AvbdOptions(iterations=5)
    scene = Scene.from_bodies(bodies, (), scenes.floor_container(20.0), opts)
    w = oracle.OracleWorld(scene, 32)
    with PhysicsContext(scene.n) as ctx:
        ctx.set_tuning(primal_lanes=lanes, fuse_dual=fuse)
        ctx.upload(scene)
        for k in range(4):
            ctx.step(opts, 15)
            w.step(opts, nsteps=15)
            _compare_world(ctx, w, 15 * (k + 1))


# ------------------------------------------------------------ edge worlds
def test_empty_world():
    """A world of no bodies uploads, steps (nothing runs) and downloads."""
    from diomedes_amd.physics import Scene

    empty = Scene.from_bodies([], (), scenes.floor_container(5.0), AvbdOptions(iterations=3))
    with PhysicsContext(4) as ctx:
        ctx.upload(empty)
        ctx.step(empty.options, 5)
        got = ctx.download()
        assert all(len(v) == 0 for v in got.values())
        assert len(ctx.contacts()) == 0 and len(ctx.pairs()) == 0
        assert ctx.stats().dynamic_bodies == 0


@pytest.mark.parametrize("mode", [0, 1])
def test_one_body_bitexact(oracle, mode):
    """A single cube dropped tilted into a container: wall contacts only, no
    pairs, 150 steps."""
    from diomedes_amd.physics import AvbdBody, Scene

    b = AvbdBody.cube((0.3, 2.0, -0.2), 0.5, 1.5)
    b.orientation = (0.2, 0.1, 0.05, float(np.sqrt(1.0 - 0.0525)))
    b.velocity = (1.0, 0.0, 0.5)
    scene = Scene.from_bodies([b], (), scenes.floor_container(3.0), AvbdOptions(iterations=6))
    w = oracle.OracleWorld(scene, 32)
    w.set_mode(mode)
    with PhysicsContext(1) as ctx:
        ctx.upload(scene)
        for k in range(3):
            ctx.step(scene.options, 50, mode=mode)
            w.step(nsteps=50)
            _compare_world(ctx, w, 50 * (k + 1))
        assert ctx.stats().pairs == 0 and ctx.stats().contacts > 0


@pytest.mark.parametrize("mode", [0, 1])
def test_static_world_bitexact(oracle, mode):
    """Only static bodies, overlapping each other and the floor: no primal
    has work, the poses never change, and the static-static contacts take
    every dual update in the final launch."""
    from diomedes_amd.physics import AvbdBody, Scene

    bodies = [AvbdBody.cube((0.7 * (i % 4), 0.4 + 0.3 * (i // 4), 0.1 * i), 0.5, 0.0) for i in range(12)]
    scene = Scene.from_bodies(bodies, (), scenes.floor_container(8.0), AvbdOptions(iterations=4))
    w = oracle.OracleWorld(scene, 32)
    w.set_mode(mode)
    with PhysicsContext(scene.n) as ctx:
        ctx.upload(scene)
        ctx.step(scene.options, 20, mode=mode)
        w.step(nsteps=20)
        _compare_world(ctx, w, 20)
        got = ctx.download()
        assert np.array_equal(got["positions"], scene.bodies["position"][:, :3])
        assert ctx.stats().dynamic_bodies == 0 and ctx.stats().contacts > 0


def test_hull_ranges_are_checked_before_any_launch(oracle):
    """A hull body names a range of the vertex pool that the kernels read
    unchecked: a record whose range is malformed (count 0, a fractional or
    negative first vertex) is rejected at upload, and one that reaches past
    the uploaded pool is rejected by the step and by the granular broad and
    narrow phases, all as DIO_INVALID before anything runs; a pool large
    enough then makes the same world step as the oracle steps it."""
    from diomedes_amd._lib import DioError
    from diomedes_amd.physics import AvbdBody, Scene, Shape, gpu_bodies

    ico = (scenes.icosphere(1) * 0.5).astype(np.float32)
    opts = AvbdOptions(iterations=4)

    def world(first, count, hull):
        bodies = [AvbdBody.cube((0.0, 0.5, 0.0), 0.5, 1.0), AvbdBody(Shape.hull(0, 42), (0.1, 1.45, 0.0)),
                  AvbdBody(Shape.hull(0, 42), (1.5, 0.6, 0.0))]
        bodies[1].inv_mass, bodies[2].inv_mass = 1.0, 1.0
        bodies[1].inv_moment = bodies[2].inv_moment = (10.0, 10.0, 10.0)
        sc = Scene.from_bodies(bodies, (), scenes.floor_container(5.0), opts)
        sc.bodies["corners"][2, 0, 0] = first
        sc.bodies["corners"][2, 0, 1] = count
        return Scene(sc.bodies, sc.state, container=sc.container, hull=hull, options=opts)

    with PhysicsContext(8, max_hull_vertices=256) as ctx:
        for first, count, what in ((0.0, 0.0, "count 0"), (0.5, 42.0, "fractional first"),
                                   (-42.0, 42.0, "negative first"), (0.0, float("nan"), "NaN count")):
            with pytest.raises(DioError) as e:
                ctx.upload(world(first, count, ico))
            assert e.value.status == "INVALID", what
        # a range past the pool: accepted at upload (the pool may follow), refused by the step
        ctx.upload(world(42.0, 42.0, ico))
        with pytest.raises(DioError) as e:
            ctx.step(opts, 1)
        assert e.value.status == "INVALID" and "84 hull vertices" in str(e.value)
        bad = world(42.0, 42.0, ico).bodies
        with pytest.raises(DioError) as e:
            ctx.broad_phase_submit(bad)
        assert e.value.status == "INVALID"
        pairs = np.zeros(1, R.GPU_PAIR)
        pairs["a"], pairs["b"] = 0, 2
        with pytest.raises(DioError) as e:
            ctx.narrow_phase_run(bad, pairs)
        assert e.value.status == "INVALID"
        # with both hulls in the pool the same world steps as the oracle's
        good = world(42.0, 42.0, np.concatenate([ico, ico]))
        ctx.upload(good)
        w = oracle.OracleWorld(good, 32)
        ctx.step(opts, 20)
        w.step(nsteps=20)
        _compare_world(ctx, w, 20)


def test_batched_table_with_empty_and_single_scenes(oracle):
    """A scene table holding empty scenes and a one-body scene between
    ordinary ones: each non-empty scene steps as the oracle steps it alone."""
    from diomedes_amd.physics import Scene

    base = scenes.batched_scenes(7, num_scenes=3, n_min=8, n_max=30).scenes
    one = base[0]
    single = Scene(one.bodies[:1].copy(), one.state[:1].copy(), container=one.container, hull=one.hull,
                   options=one.options)
    none = Scene(one.bodies[:0].copy(), one.state[:0].copy(), container=one.container, hull=one.hull,
                 options=one.options)
    parts = [none, base[0], none, single, base[1], none, base[2], none]
    batch = scenes.BatchedScenes(parts, one.hull, one.options)
    worlds = {s: oracle.OracleWorld(sc, 32) for s, sc in enumerate(parts) if sc.n > 0}
    with PhysicsContext(batch.n, max_hull_vertices=len(batch.hull)) as ctx:
        ctx.batch_upload(batch)
        for k in range(2):
            ctx.step(batch.options, 25)
            got = ctx.download()
            for s, w in worlds.items():
                w.step(nsteps=25)
                ref = w.state()
                sl = batch.scene_slice(s)
                for key in ("positions", "orientations", "velocities", "angular_velocities"):
                    assert_bits(got[key][sl], ref[key], f"scene {s} step {25 * (k + 1)} {key}")


# ------------------------------------------------------------ batched world
def _scene_contacts(gc, lo, hi):
    """The contacts of bodies lo .. hi - 1 in a batched contact list, with
    scene-local indices (the container stays CONTAINER)."""
    sel = (gc["a"] >= lo) & (gc["a"] < hi)
    c = gc[sel].copy()
    c["a"] -= lo
    notc = c["b"] != 0xFFFFFFFF
    c["b"][notc] -= lo
    return c


@pytest.mark.parametrize("tuning", [{}, {"primal_packed": 2, "fuse_dual": 1}, {"primal_packed": 2, "fuse_dual": 0}])
def test_batched_scenes_match_oracle_per_scene(oracle, tuning):
    """Every scene of a batched world steps bit-identically to the FP32 oracle
    stepping that scene alone (boxes and icosphere hulls, own floors); with the
    library's default primal and with the span primal."""
    batch = scenes.batched_scenes(3, num_scenes=6, n_min=8, n_max=48)
    opts = batch.options
    worlds = [oracle.OracleWorld(sc, 32) for sc in batch.scenes]
    with PhysicsContext(batch.n, max_hull_vertices=len(batch.hull)) as ctx:
        if tuning:
            ctx.set_tuning(**tuning)
        ctx.batch_upload(batch)
        for k in range(3):
            ctx.step(opts, 20)
            got = ctx.download()
            gc = ctx.contacts()
            for s, w in enumerate(worlds):
                w.step(nsteps=20)
                sl = batch.scene_slice(s)
                ref = w.state()
                for key in ("positions", "orientations", "velocities", "angular_velocities", "prev_velocities"):
                    assert_bits(got[key][sl], ref[key], f"scene {s} step {20 * (k + 1)} {key}")
                c = _scene_contacts(gc, sl.start, sl.stop)
                rc = w.contacts()
                assert len(c) == len(rc["a"]), (s, len(c), len(rc["a"]))
                assert np.array_equal(c["a"], rc["a"]) and np.array_equal(c["b"], rc["b"])
                assert_bits(c["lambda_"][:, :3], rc["lambda"], f"scene {s} lambda")
            _print("batched step", 20 * (k + 1), "contacts", len(gc))


def _step_batch(batch, steps, **tuning):
    with PhysicsContext(batch.n, max_pairs=8 * batch.n, max_contacts=16 * batch.n,
                        max_hull_vertices=max(len(batch.hull), 1)) as ctx:
        ctx.set_tuning(**tuning)
        ctx.batch_upload(batch)
        out = []
        for chunk in steps:
            ctx.step(batch.options, chunk)
            out.append((ctx.download(), ctx.contacts()))
        return out


@pytest.mark.parametrize("bundle", [1, 300])
@pytest.mark.parametrize("fuse", [0, 1])
@pytest.mark.parametrize("table", ["mixed", "many_small", "with_empty"])
def test_scene_solve_bitexact(table, fuse, bundle):
    """scene_solve N (the whole solve in one launch, each workgroup taking
    bundles of consecutive scenes of at least N bodies from a queue) against
    the per-launch schedule: every body and contact bit-identical, over
    several steps (the queue is reset by each launch's last workgroup), for a
    table wider than the launch (3000 one-scene bundles against 2048
    workgroups) and one holding empty and one-body scenes."""
    from diomedes_amd.physics import Scene

    if table == "mixed":
        batch = scenes.batched_scenes(5, num_scenes=24, n_min=8, n_max=512)
    elif table == "many_small":
        batch = scenes.batched_scenes(9, num_scenes=3000, n_min=8, n_max=24)
    else:
        base = scenes.batched_scenes(7, num_scenes=3, n_min=8, n_max=30).scenes
        one = base[0]
        single = Scene(one.bodies[:1].copy(), one.state[:1].copy(), container=one.container, hull=one.hull,
                       options=one.options)
        none = Scene(one.bodies[:0].copy(), one.state[:0].copy(), container=one.container, hull=one.hull,
                     options=one.options)
        batch = scenes.BatchedScenes([none, base[0], none, single, base[1], none, base[2], none], one.hull,
                                     one.options)
    steps = (1, 14, 30)
    ref = _step_batch(batch, steps, primal_lanes=1, fuse_dual=fuse, scene_solve=0)
    got = _step_batch(batch, steps, primal_lanes=1, fuse_dual=fuse, scene_solve=bundle)
    done = 0
    for (g, gc), (r, rc), k in zip(got, ref, steps):
        done += k
        for key in ("positions", "orientations", "velocities", "angular_velocities", "prev_velocities"):
            assert_bits(g[key], r[key], f"{table} fuse {fuse} step {done} {key}")
        assert len(gc) == len(rc)
        for key in ("lambda_", "penalty"):
            assert_bits(gc[key], rc[key], f"{table} fuse {fuse} step {done} contact {key}")
    _print(f"scene solve {table} ({len(batch.scenes)} scenes, {batch.n} bodies) fuse {fuse} bundle {bundle}: "
           "bit-identical")


def test_batched_scenes_match_plain_contexts():
    """A 24-scene batch (up to 512 bodies per scene) against each scene stepped
    in a plain context of its own, GPU against GPU."""
    batch = scenes.batched_scenes(5, num_scenes=24, n_min=8, n_max=512)
    opts = batch.options
    with PhysicsContext(batch.n, max_hull_vertices=len(batch.hull)) as ctx:
        ctx.batch_upload(batch)
        ctx.step(opts, 45)
        got = ctx.download()
    for s, sc in enumerate(batch.scenes):
        with PhysicsContext(sc.n, max_hull_vertices=len(sc.hull)) as one:
            one.upload(sc)
            one.step(opts, 45)
            ref = one.download()
        sl = batch.scene_slice(s)
        for key in ("positions", "orientations", "velocities", "angular_velocities"):
            assert_bits(got[key][sl], ref[key], f"scene {s} ({sc.n} bodies) {key}")

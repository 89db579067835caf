"""GPU hard joints (kinds 4-6) against the FP32 oracle: bit-identical poses,
velocities and multipliers (lambda and penalty per row), in the coloured mode
(joints between dynamic bodies) and in the Jacobi mode's fused and unfused
schedules (joints to static anchors), plus the errors the world path gives.
"""

from __future__ import annotations

import numpy as np
import pytest

from diomedes_amd import PhysicsContext, records as R, scenes
from diomedes_amd._lib import DioError
from diomedes_amd.physics import AvbdBody, AvbdOptions, HardBallSocket, HardRod, HardRope, Scene

from .helpers import assert_bits, primal_tuning

pytestmark = pytest.mark.gpu

STATE_KEYS = ("positions", "orientations", "velocities", "angular_velocities")


def _print(*a):
    print(*a, flush=True)


def _compare(ctx, w, what):
    g, o = ctx.download(), w.state()
    for key in STATE_KEYS:
        assert_bits(g[key], o[key], f"{what}: {key}")
    gl, gp = ctx.joints()
    ol, op = w.joints()
    assert_bits(gl, ol, f"{what}: joint lambda")
    assert_bits(gp, op, f"{what}: joint penalty")


def _pendulums(count=48):
    """Pendulums on hard joints to static anchors (no joint between two
    dynamic bodies): rods, off-centre sockets and ropes, released sideways
    into one another."""
    bodies = []
    cons = []
    for i in range(count):
        ax, az = (i % 8) * 1.5, (i // 8) * 1.5
        bodies.append(AvbdBody.sphere((ax, 6.0, az), 0.1, 0.0))
        a = len(bodies) - 1
        bodies.append(AvbdBody.cube((ax + 0.9 + 0.05 * (i % 5), 5.6 - 0.03 * (i % 7), az + 0.3), 0.2,
                                    0.5 + 0.1 * (i % 4)))
        b = len(bodies) - 1
        if i % 3 == 0:
            cons.append(HardRod(a, b, (0.0, 0.0, 0.0), (0.05, 0.1, 0.0), 1.1 + 0.02 * (i % 6)))
        elif i % 3 == 1:
            cons.append(HardBallSocket(b, a, (0.0, 0.2, 0.05), (0.95, -0.2, 0.3)))
        else:
            cons.append(HardRope(a, b, (0.0, 0.0, 0.0), (0.0, 0.2, 0.0), 1.3))
    return Scene.from_bodies(bodies, cons, scenes.floor_container(30.0), AvbdOptions(iterations=6))


@pytest.mark.parametrize("lanes", [32, 8, 1])
def test_hard_chains_coloured_bitexact(oracle, lanes):
    sc = scenes.hard_chains(1, chains=12, links=10)
    w = oracle.OracleWorld(sc, 32)
    w.set_mode(1)
    with PhysicsContext(sc.n, max_constraints=len(sc.constraints)) as ctx:
        ctx.set_tuning(primal_lanes=lanes)
        ctx.upload(sc)
        for k in range(4):
            ctx.step(sc.options, 15, mode=1)
            w.step(nsteps=15)
            _compare(ctx, w, f"lanes {lanes} step {15 * (k + 1)}")
            st = ctx.stats()
            _print("hard chains lanes", lanes, "step", 15 * (k + 1), "contacts", st.contacts, "colours", st.colours)
        assert st.contacts > 50
        lam, pen = ctx.joints()
        assert np.all(pen.max(axis=1) > 1.0)


@pytest.mark.parametrize("lanes,fuse", [(32, 1), (8, 1), (1, 1), (32, 0), ("packed", 1), ("packed", 0)])
def test_pendulums_jacobi_bitexact(oracle, lanes, fuse):
    """Jacobi mode: the fused schedule (each joint's dual update inside the
    next primal, written by the dynamic body) and the unfused one."""
    sc = _pendulums()
    w = oracle.OracleWorld(sc, 32)
    with PhysicsContext(sc.n, max_constraints=len(sc.constraints)) as ctx:
        ctx.set_tuning(**primal_tuning(lanes), fuse_dual=fuse)
        ctx.upload(sc)
        for k in range(3):
            ctx.step(sc.options, 20, mode=0)
            w.step(nsteps=20)
            _compare(ctx, w, f"lanes {lanes} fuse {fuse} step {20 * (k + 1)}")
        _print("pendulums lanes", lanes, "fuse", fuse, "contacts", ctx.stats().contacts)


def test_options_change_between_calls(oracle):
    """The joints' warm start follows the options of the step that uses it
    (alpha, gamma and the penalty range change between calls)."""
    sc = _pendulums(24)
    w = oracle.OracleWorld(sc, 32)
    o1 = sc.options
    o2 = AvbdOptions(iterations=6, alpha=0.95, gamma=0.99, penalty_min=10.0, penalty_max=1.0e8)
    with PhysicsContext(sc.n, max_constraints=len(sc.constraints)) as ctx:
        ctx.upload(sc)
        for k, o in enumerate((o1, o2, o2, o1, o2)):
            ctx.step(o, 7)
            w.step(o, nsteps=7)
            _compare(ctx, w, f"phase {k}")


def test_modes_alternate_with_hard_joints(oracle):
    sc = _pendulums(24)
    w = oracle.OracleWorld(sc, 32)
    with PhysicsContext(sc.n, max_constraints=len(sc.constraints)) as ctx:
        ctx.upload(sc)
        for k, mode in enumerate((0, 1, 1, 0)):
            w.set_mode(mode)
            ctx.step(sc.options, 9, mode=mode)
            w.step(nsteps=9)
            _compare(ctx, w, f"phase {k} mode {mode}")


class _Table:
    """Several scenes in one batched world (constraints in global indices)."""

    def __init__(self, parts):
        self.bodies = np.concatenate([p.bodies for p in parts])
        self.state = np.concatenate([p.state for p in parts])
        offs = np.cumsum([0] + [p.n for p in parts]).astype(np.uint32)
        cons = []
        for p, o in zip(parts, offs[:-1]):
            c = p.constraints.copy()
            c["a"] += o
            c["b"] += o
            cons.append(c)
        self.constraints = np.concatenate(cons)
        self.offsets = offs
        self.containers = [p.container for p in parts]
        self.hull = np.zeros((0, 3), np.float32)
        self.n = int(offs[-1])


def test_batched_hard_chains_match_plain():
    parts = [scenes.hard_chains(s, chains=4 + s, links=6 + 2 * s) for s in range(4)]
    table = _Table(parts)
    opts = parts[0].options
    with PhysicsContext(table.n, max_constraints=len(table.constraints)) as ctx:
        ctx.batch_upload(table)
        ctx.step(opts, 30, mode=1)
        gb = ctx.download()
        lb, pb = ctx.joints()
    for s, sc in enumerate(parts):
        with PhysicsContext(sc.n, max_constraints=len(sc.constraints)) as one:
            one.upload(sc)
            one.step(opts, 30, mode=1)
            g1 = one.download()
            l1, p1 = one.joints()
        sl = slice(int(table.offsets[s]), int(table.offsets[s + 1]))
        c0 = sum(len(p.constraints) for p in parts[:s])
        cs = slice(c0, c0 + len(sc.constraints))
        for key in STATE_KEYS:
            assert_bits(gb[key][sl], g1[key], f"scene {s}: {key}")
        assert_bits(lb[cs], l1, f"scene {s}: lambda")
        assert_bits(pb[cs], p1, f"scene {s}: penalty")


def test_hard_chains_full_size_bitexact(oracle):
    """256 chains of 32 links (8,448 bodies, 8,192 hard joints) in the
    coloured mode, 90 steps: the chains swing down, hit the floor and pile
    into each other."""
    sc = scenes.hard_chains(0)
    w = oracle.OracleWorld(sc, 32)
    w.set_mode(1)
    with PhysicsContext(sc.n, max_constraints=len(sc.constraints)) as ctx:
        ctx.upload(sc)
        for k, nsteps in enumerate((40, 30, 20)):
            ctx.step(sc.options, nsteps, mode=1)
            w.step(nsteps=nsteps)
            _compare(ctx, w, f"block {k}")
            st = ctx.stats()
            _print("hard chains 256x32 block", k, "contacts", st.contacts, "colours", st.colours)
        assert st.contacts > 1000


@pytest.mark.parametrize("bend,lanes", [("soft", 32), ("hard", 32), ("hard", 1)])
def test_bending_chains_bitexact(oracle, bend, lanes):
    """Bending joints beside the socket chains' hard sockets (kinds 7 and 8),
    coloured mode: poses, velocities and multipliers bit-identical."""
    sc = scenes.hard_chains(3, chains=12, links=10, bend=bend)
    w = oracle.OracleWorld(sc, 32)
    w.set_mode(1)
    with PhysicsContext(sc.n, max_constraints=len(sc.constraints)) as ctx:
        ctx.set_tuning(primal_lanes=lanes)
        ctx.upload(sc)
        for k in range(3):
            ctx.step(sc.options, 20, mode=1)
            w.step(nsteps=20)
            _compare(ctx, w, f"{bend} bend lanes {lanes} step {20 * (k + 1)}")
        _print(bend, "bend chains lanes", lanes, "contacts", ctx.stats().contacts, "colours", ctx.stats().colours)


@pytest.mark.parametrize("lanes,fuse", [(32, 1), (1, 1), (32, 0), ("packed", 1), ("packed", 0)])
def test_cantilevers_jacobi_bitexact(oracle, lanes, fuse):
    """Single links held from static anchors by a hard socket and a hard bend
    (no joint between dynamic bodies: the Jacobi mode's fused and unfused
    schedules)."""
    sc = scenes.hard_chains(4, chains=12, links=1, bend="hard")
    w = oracle.OracleWorld(sc, 32)
    with PhysicsContext(sc.n, max_constraints=len(sc.constraints)) as ctx:
        ctx.set_tuning(**primal_tuning(lanes), fuse_dual=fuse)
        ctx.upload(sc)
        for k in range(3):
            ctx.step(sc.options, 20, mode=0)
            w.step(nsteps=20)
            _compare(ctx, w, f"cantilevers lanes {lanes} fuse {fuse} step {20 * (k + 1)}")


def test_hard_joint_errors():
    sc = scenes.hard_chains(2, chains=3, links=4)
    with PhysicsContext(sc.n, max_constraints=len(sc.constraints)) as ctx:
        ctx.upload(sc)
        with pytest.raises(DioError, match="coloured mode"):
            ctx.step(sc.options, 1, mode=0)
        ctx.step(sc.options, 0, mode=0)  # no step runs: accepted
        ctx.step(sc.options, 2, mode=1)
        with pytest.raises(DioError, match="joints"):
            ctx.adjoint_begin(4, 2)
    # the granular solve (dio_avbd_submit) carries no multiplier state
    bodies = [AvbdBody.sphere((0.0, 2.0, 0.0), 0.1, 0.0), AvbdBody.sphere((0.0, 1.0, 0.0), 0.2, 1.0)]
    one = Scene.from_bodies(bodies, [HardRod(0, 1, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 1.0)])
    with PhysicsContext(2, max_constraints=1) as ctx:
        with pytest.raises(DioError, match="hard joints"):
            ctx.gpu_physics_submit(one.state, one.bodies["position"][:, :3], one.bodies["orientation"],
                                   np.zeros(0, R.AVBD_CONTACT), one.constraints, np.zeros(3, np.uint32),
                                   np.zeros(0, np.uint32), AvbdOptions(iterations=2))

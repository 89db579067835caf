"""The oracle's hard joints (kinds 4-6: hard rod, socket and rope; no
counterpart in the reference, whose joints are soft penalties).

Pinned by what the mathematics fixes rather than by the oracle's own
formulas:
- the first dual update of a body hanging at rest at the joint's length,
  derived by hand: one primal step along y, then lambda = C^ and the penalty
  ramp, in closed form;
- equilibrium: a hanging body's multiplier converges to its weight m |g|
  (rod, socket, rope);
- a slack rope exerts nothing: the body's trajectory is free fall, bit for
  bit, until the rope is taut, after which it hangs at the rope's length;
- pendulums on static anchors share no joint between dynamic bodies, so the
  coloured Gauss-Seidel solve is the Jacobi solve, bit for bit;
- a chain on hard rods keeps its link lengths far better than on the
  reference's soft rods, in FP64 and FP32 alike;
- bending joints (kinds 7, 8): a body on a soft bend is a linear torsional
  oscillator whose angle follows backward Euler's recurrence
  th_{n+1} = (2 th_n - th_{n-1}) / (1 + w^2 dt^2), w^2 = k / I; a cantilever of
  N links on hard sockets and hard bends carries, at joint j, the shear
  m g (N - j) and the moment m g s (N - j)^2 / 2; a bend moves no body.
"""

from __future__ import annotations

import numpy as np
import pytest

from diomedes_amd import scenes
from diomedes_amd.physics import AvbdBody, AvbdOptions, Bend, HardBallSocket, HardRod, HardRope, Rod, Scene

G = 10.0
DT = float(np.float32(1.0 / 60.0))  # the options record holds dt in float32


def _hanging(kind, length=1.0, mass=2.0, offset=0.0, iterations=8, start=None):
    """A body below a static anchor at (0, 10, 0), held by one hard joint."""
    y = 10.0 - length if start is None else start
    bodies = [AvbdBody.sphere((0.0, 10.0, 0.0), 0.1, 0.0), AvbdBody.sphere((offset, y, 0.0), 0.2, mass)]
    if kind == "rod":
        c = HardRod(0, 1, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), length)
    elif kind == "socket":
        c = HardBallSocket(0, 1, (0.0, 0.0, 0.0), (0.0, length, 0.0))
    else:
        c = HardRope(0, 1, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), length)
    return Scene.from_bodies(bodies, [c], None, AvbdOptions(iterations=iterations))


@pytest.mark.parametrize("kind", ["rod", "socket", "rope"])
def test_first_update_closed_form(oracle, kind):
    """At rest at the joint's length, one iteration: the primal moves the body
    by dy = -m g / (m / dt^2 + k) with k = penalty_min = 1 (the joint's row
    has zero force, C^ = 0, and adds k to H_yy; a rope's row with zero force
    is slack and adds nothing, k = 0); the dual then sets lambda = 1 * C^ =
    -dy and ramps the penalty to 1 + beta |dy|."""
    m = 2.0
    sc = _hanging(kind, mass=m, iterations=1)
    w = oracle.OracleWorld(sc, 64)
    w.step(nsteps=1)
    dy = -m * G / (m / DT**2 + (0.0 if kind == "rope" else 1.0))
    y = w.state()["positions"][1, 1]
    assert abs(y - (9.0 + dy)) < 1e-12
    lam, pen = w.joints()
    row = 1 if kind == "socket" else 0
    assert lam[0, row] == pytest.approx(-dy, rel=1e-9)
    assert pen[0, row] == pytest.approx(1.0 + 1.0e4 * abs(dy), rel=1e-9)
    if kind == "socket":
        # the x and z rows stay at rest: no force, no ramp
        assert np.all(lam[0, [0, 2]] == 0.0) and np.all(pen[0, [0, 2]] == 1.0)


@pytest.mark.parametrize("kind", ["rod", "socket", "rope"])
def test_equilibrium_multiplier_is_the_weight(oracle, kind):
    m = 2.0
    w = oracle.OracleWorld(_hanging(kind, mass=m), 64)
    w.step(nsteps=600)
    lam, _ = w.joints()
    row = 1 if kind == "socket" else 0
    assert lam[0, row] == pytest.approx(m * G, rel=2e-3)
    st = w.state()
    assert abs(st["positions"][1, 1] - 9.0) < 2e-3
    assert np.max(np.abs(st["velocities"][1])) < 1e-3


@pytest.mark.parametrize("precision", [32, 64])
def test_slack_rope_is_free_fall(oracle, precision):
    rope = _hanging("rope", length=2.0, mass=1.0, start=9.5)
    free = Scene.from_bodies([AvbdBody.sphere((0.0, 10.0, 0.0), 0.1, 0.0), AvbdBody.sphere((0.0, 9.5, 0.0), 0.2, 1.0)],
                             [], None, AvbdOptions(iterations=8))
    wr, wf = oracle.OracleWorld(rope, precision), oracle.OracleWorld(free, precision)
    slack_steps = 0
    for _ in range(30):
        wr.step(nsteps=1)
        wf.step(nsteps=1)
        a, b = wr.state(), wf.state()
        if 10.0 - b["positions"][1, 1] >= 2.0:
            break
        slack_steps += 1
        for k in ("positions", "orientations", "velocities"):
            assert np.array_equal(a[k], b[k]), k
        lam, _ = wr.joints()
        assert lam[0, 0] == 0.0
    assert slack_steps >= 25
    wr.step(nsteps=300)
    st = wr.state()
    assert abs(st["positions"][1, 1] - 8.0) < 2e-3
    lam, _ = wr.joints()
    assert lam[0, 0] == pytest.approx(G, rel=5e-3)


def _pendulums():
    bodies = [AvbdBody.sphere((0.0, 10.0, 0.0), 0.1, 0.0)]
    cons = []
    for i in range(5):
        # each pendulum released sideways, off-centre anchors on the body so the
        # rows couple rotation
        bodies.append(AvbdBody.cube((1.0 + 0.3 * i, 9.5 - 0.1 * i, 0.2 * i), 0.1, 0.5 + 0.25 * i))
        k = len(bodies) - 1
        if i % 3 == 0:
            cons.append(HardRod(0, k, (0.0, 0.0, 0.0), (0.05, 0.1, 0.0), 1.2 + 0.1 * i))
        elif i % 3 == 1:
            cons.append(HardBallSocket(k, 0, (0.0, 0.1, 0.05), (1.0 + 0.3 * i, -0.4, 0.2 * i)))
        else:
            cons.append(HardRope(0, k, (0.0, 0.0, 0.0), (0.0, 0.1, 0.0), 1.0 + 0.3 * i))
    return Scene.from_bodies(bodies, cons, None, AvbdOptions(iterations=6))


@pytest.mark.parametrize("precision", [32, 64])
def test_jacobi_equals_gauss_seidel_without_dynamic_pairs(oracle, precision):
    sc = _pendulums()
    out = []
    for mode in (0, 1):
        w = oracle.OracleWorld(sc, precision)
        w.set_mode(mode)
        w.step(nsteps=40)
        out.append((w.state(), w.joints()))
    for k in ("positions", "orientations", "velocities", "angular_velocities"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert np.array_equal(out[0][1][0], out[1][1][0]) and np.array_equal(out[0][1][1], out[1][1][1])
    assert np.all(out[0][1][1].max(axis=1) > 1.0)  # every joint became active and ramped


def _chain(hard, links=10, length=0.5):
    bodies = [AvbdBody.sphere((0.0, 10.0, 0.0), 0.1, 0.0)]
    cons = []
    for i in range(links):
        bodies.append(AvbdBody.cube((length * (i + 1), 10.0, 0.0), 0.1, 1.0))
        cons.append((HardRod if hard else Rod)(i, i + 1, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), length))
    return Scene.from_bodies(bodies, cons, None, AvbdOptions(iterations=8))


def _max_length_error(sc, st):
    p = st["positions"]
    return max(abs(np.linalg.norm(p[c["a"]] - p[c["b"]]) - c["anchor_a"][3]) for c in sc.constraints)


def test_hard_chain_holds_its_length(oracle):
    """A chain released horizontally swings down (coloured mode): over 200
    steps its worst link error on hard rods stays far below the soft rods'."""
    worst = {}
    final = {}
    for hard in (False, True):
        for prec in (64, 32):
            sc = _chain(hard)
            w = oracle.OracleWorld(sc, prec)
            w.set_mode(1)
            errs = []
            for _ in range(20):
                w.step(nsteps=10)
                errs.append(_max_length_error(sc, w.state()))
            worst[hard, prec] = max(errs[5:])
            final[hard, prec] = w.state()["positions"]
    assert worst[True, 64] < worst[False, 64] / 5
    assert worst[True, 64] < 0.03
    # FP32 follows FP64 (a smooth swing: no contacts)
    assert np.max(np.abs(final[True, 32] - final[True, 64])) < 1e-2
    assert worst[True, 32] < 0.03


def test_soft_bend_is_a_backward_euler_oscillator(oracle):
    """A cube on a soft bend to a static body, no gravity, started at a
    small angle about z: one body, a quadratic energy in the angle, so the
    step is backward Euler on I th'' = -k th."""
    th0, k = 0.04, 1.0
    bodies = [AvbdBody.sphere((5.0, 0.0, 0.0), 0.1, 0.0), AvbdBody.cube((0.0, 0.0, 0.0), 0.1, 1.0)]
    bodies[1].orientation = (0.0, 0.0, float(np.sin(th0 / 2)), float(np.cos(th0 / 2)))
    sc = Scene.from_bodies(bodies, [Bend(0, 1, stiffness=k)], None, AvbdOptions(iterations=4, gravity=(0.0, 0.0, 0.0)))
    w2dt2 = k / (1.0 / float(sc.state["inv_moment"][1][0])) * DT * DT
    w = oracle.OracleWorld(sc, 64)
    th = [th0]
    for _ in range(60):
        w.step(nsteps=1)
        q = w.state()["orientations"][1]
        th.append(2.0 * np.arctan2(q[2], q[3]))
        assert np.all(w.state()["positions"][1] == 0.0)  # the bend's rows are angular only
    ref = [th0, th0 / (1.0 + w2dt2)]
    for _ in range(59):
        ref.append((2.0 * ref[-1] - ref[-2]) / (1.0 + w2dt2))
    # the error is the bend error's own nonlinearity, 2 sin(th / 2) against th
    assert np.max(np.abs(np.array(th) - np.array(ref))) < 2e-4 * th0
    assert min(th) < -0.5 * th0  # it swings through


def test_cantilever_shear_and_moment(oracle):
    """Six unit-mass links on hard sockets and hard bends, held horizontally
    from a static anchor: at equilibrium joint j carries the weight of the
    links beyond it and their moment about it."""
    n_links, s = 6, 0.34
    sc = scenes.hard_chains(0, chains=1, links=n_links, bend="hard")
    w = oracle.OracleWorld(sc, 64)
    w.set_mode(1)
    w.step(nsteps=600)
    lam, _ = w.joints()
    kinds = sc.constraints["kind"]
    g = 9.81
    shear = np.abs(lam[kinds == 5][:, 1])
    moment = np.abs(lam[kinds == 8][:, 2])
    for j in range(n_links):
        assert shear[j] == pytest.approx(g * (n_links - j), rel=2e-3), j
        assert moment[j] == pytest.approx(g * s * (n_links - j) ** 2 / 2.0, rel=2e-3), j
    assert np.max(np.abs(w.state()["velocities"])) < 1e-3


"""The FP64 oracle gradient of the frozen-contact replay.

The adjoint differentiates a replay in which each step's contact set
(geometry and warm-start multipliers, as the solve received them) is frozen
and only the AVBD solve runs.  The oracle replays a recording in float32 or
float64, or in float64 with forward-mode dual numbers.  These tests pin the
replay (it reproduces the recorded run bit for bit) and the dual-number
derivatives (central differences of the float64 replay).
"""

from __future__ import annotations

import numpy as np

from diomedes_amd import scenes


def _loss(pos, target, nb):
    d = pos[:nb] - target
    return float(np.mean(np.sum(d * d, axis=1)))


def _throw(steps=40):
    sc = scenes.wall_throw(1, columns=6, rows=4)
    return sc, sc.n - 1


def test_replay_reproduces_the_recorded_run(oracle):
    sc, _ = _throw()
    rec = oracle.record_world(sc, 40, precision=32)
    w = oracle.OracleWorld(sc, 32)
    w.step(nsteps=40)
    ref = w.state()
    pos, rot = oracle.replay(sc.bodies, sc.state, rec, sc.options, sc.options.friction, precision=32)
    assert np.array_equal(pos.astype(np.float32), ref["positions"].astype(np.float32))
    assert np.array_equal(rot.astype(np.float32), ref["orientations"].astype(np.float32))


def _ad_and_fd(oracle, sc, proj, steps, h_rel=1e-5):
    nb = sc.n - 1
    rec = oracle.record_world(sc, steps, precision=32)
    target = scenes.wall_throw_target(sc, 1)
    p0 = np.array([*sc.state["vel"][proj, :3].astype(np.float64), 1.0 / float(sc.state["inv_mass"][proj]),
                   float(np.float32(sc.options.friction))])
    pos, _, dpos, _ = oracle.replay(sc.bodies, sc.state, rec, sc.options, p0[4], precision=65, seed_body=proj,
                                    params=p0)
    grad_ad = np.einsum("ij,ijk->k", 2.0 * (pos[:nb] - target) / nb, dpos[:nb])
    grad_fd = np.zeros(5)
    for k in range(5):
        h = h_rel * max(abs(p0[k]), 1.0)
        lp = _loss(oracle.replay(sc.bodies, sc.state, rec, sc.options, 0.0, 64, proj, p0 + h * np.eye(5)[k])[0],
                   target, nb)
        lm = _loss(oracle.replay(sc.bodies, sc.state, rec, sc.options, 0.0, 64, proj, p0 - h * np.eye(5)[k])[0],
                   target, nb)
        grad_fd[k] = (lp - lm) / (2 * h)
    return grad_ad, grad_fd


def test_dual_number_gradient_matches_central_differences(oracle):
    """d loss / d (v0, m, mu) of the float64 replay: forward-mode dual numbers
    against central differences, where the replayed function is smooth at the
    difference step: v0 and mass over 20 steps (the sphere has hit the wall),
    mu over 10 (later, contacts switch between the friction cone's pieces
    within the difference step)."""
    sc = scenes.wall_throw(1, columns=6, rows=8)
    proj = sc.n - 1
    ad20, fd20 = _ad_and_fd(oracle, sc, proj, 20)
    ad10, fd10 = _ad_and_fd(oracle, sc, proj, 10)
    print("20 steps: dual numbers", ad20, "central differences", fd20)
    print("10 steps: dual numbers", ad10, "central differences", fd10)
    for k in range(4):
        assert abs(ad20[k] - fd20[k]) <= 1e-5 * np.max(np.abs(fd20[:4])), (k, ad20, fd20)
    assert abs(ad10[4] - fd10[4]) <= 1e-5 * abs(fd10[4]), (ad10, fd10)
    assert np.all(np.abs(ad20[:4]) > 0) and abs(ad10[4]) > 0


def test_float_valued_dual_numbers_follow_the_float_replay(oracle):
    """precision 33: the values are the float replay's, bit for bit, and the
    derivatives agree with the float64 dual numbers where the replay is
    smooth (20 steps, v0 and mass)."""
    sc = scenes.wall_throw(1, columns=6, rows=8)
    proj = sc.n - 1
    rec = oracle.record_world(sc, 20, precision=32)
    p0 = np.array([*sc.state["vel"][proj, :3].astype(np.float64), 1.0 / float(sc.state["inv_mass"][proj]),
                   float(np.float32(sc.options.friction))])
    pos32, rot32 = oracle.replay(sc.bodies, sc.state, rec, sc.options, p0[4], 32, proj, p0)
    pos33, rot33, dpos33, _ = oracle.replay(sc.bodies, sc.state, rec, sc.options, p0[4], 33, proj, p0)
    _, _, dpos65, _ = oracle.replay(sc.bodies, sc.state, rec, sc.options, p0[4], 65, proj, p0)
    assert np.array_equal(pos32, pos33) and np.array_equal(rot32, rot33)
    scale = np.max(np.abs(dpos65[:, :, :4]))
    assert scale > 0
    assert np.max(np.abs(dpos33[:, :, :4] - dpos65[:, :, :4])) <= 1e-3 * scale

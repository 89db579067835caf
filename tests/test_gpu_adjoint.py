"""GPU adjoint: the recording, the frozen-contact replay and the hand-written
reverse mode, through the C ABI.

- The replay at the recorded parameters reproduces the recorded run bit for
  bit, and the replay from other velocities / masses / friction equals the
  float32 oracle replay of the same recording bit for bit.
- The reverse-mode gradient of a loss on the final poses with respect to the
  projectile's initial velocity and mass and the friction coefficient is
  checked against the oracle's forward-mode derivatives along the float
  replay (dual numbers with float values and float64 tangents: the same
  branches as the GPU run), and against the float64 dual numbers and central
  differences of the float64 replay where that replay is smooth.
- The frozen-multiplier mode (the multipliers and penalties each dual update
  writes held constant) against the oracle's frozen replay: its float-valued
  dual numbers, and float64 central differences at h = 1e-3 on the 64 x 32
  wall; over the full 120-step throw it stays finite.
"""

from __future__ import annotations

import numpy as np
import pytest

from diomedes_amd import PhysicsContext, scenes
from diomedes_amd import records as R

from .helpers import assert_bits

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _throw(rows=8, columns=6, seed=1):
    sc = scenes.wall_throw(seed, columns=columns, rows=rows)
    return sc, sc.n - 1


def _record(sc, steps, every=8, **tuning):
    ctx = PhysicsContext(sc.n)
    if tuning:
        ctx.set_tuning(**tuning)
    ctx.upload(sc)
    ctx.adjoint_begin(steps, every)
    ctx.step(sc.options, steps)
    return ctx


def _params(sc, proj):
    return np.array([*sc.state["vel"][proj, :3].astype(np.float64), 1.0 / float(sc.state["inv_mass"][proj]),
                     float(np.float32(sc.options.friction))])


def _state_at(sc, proj, p):
    """The body state with the projectile's v0 / mass set as the oracle's
    parameterised replay sets them (float32 arithmetic)."""
    st = sc.state.copy()
    st["vel"][proj, :3] = np.float32(p[:3])
    inv_m = float(sc.state["inv_mass"][proj])
    c = sc.state["inv_moment"][proj, :3].astype(np.float64) / inv_m
    m = np.float32(p[3])
    st["inv_mass"][proj] = np.float32(1.0) / m
    st["inv_moment"][proj, :3] = c.astype(np.float32) / m
    return st


def _options(sc, mu):
    import dataclasses

    return dataclasses.replace(sc.options, friction=float(np.float32(mu)))


def _gpu_grad(ctx, sc, proj, p, target, state0=None):
    """d loss / d (v0 x, y, z, m, mu) from the GPU reverse mode, with
    loss = mean over bricks of |x_final - target|^2."""
    nb = sc.n - 1
    opts = _options(sc, p[4])
    pos, _ = ctx.adjoint_replay(opts, state0)
    dpos = np.zeros((sc.n, 3), np.float32)
    dpos[:nb] = (2.0 * (pos[:nb].astype(np.float64) - target) / nb).astype(np.float32)
    g = ctx.adjoint_backward(opts, d_positions=dpos, state0=state0)
    inv_m = float(sc.state["inv_mass"][proj])
    c = sc.state["inv_moment"][proj, :3].astype(np.float64) / inv_m
    m = p[3]
    gm = -(float(g["inv_mass"][proj]) + float(np.dot(g["inv_moment"][proj].astype(np.float64), c))) / (m * m)
    return np.array([*g["velocities"][proj].astype(np.float64), gm, g["friction"]]), pos, g


def _oracle_grad(oracle, sc, proj, rec, p, target, precision):
    nb = sc.n - 1
    pos, _, dpos, _ = oracle.replay(sc.bodies, sc.state, rec, sc.options, p[4], precision=precision, seed_body=proj,
                                    params=p)
    return np.einsum("ij,ijk->k", 2.0 * (pos[:nb] - target) / nb, dpos[:nb]), pos


# ------------------------------------------------------------------ replay
@pytest.mark.parametrize("lanes,fuse", [(32, 1), (1, 0)])
def test_replay_reproduces_the_recorded_run(lanes, fuse):
    sc, _ = _throw()
    steps = 20
    ctx = _record(sc, steps, every=8, primal_lanes=lanes, fuse_dual=fuse)
    assert ctx.adjoint_steps() == steps
    ref = ctx.download()
    for state0 in (None, sc.state):
        pos, rot = ctx.adjoint_replay(sc.options, state0)
        assert_bits(pos, ref["positions"], "replay positions")
        assert_bits(rot, ref["orientations"], "replay orientations")
    _print("replay reproduces", steps, "steps at lanes", lanes, "fuse", fuse)
    ctx.close()


def test_replay_from_other_parameters_matches_the_oracle(oracle):
    sc, proj = _throw()
    steps = 24
    rec = oracle.record_world(sc, steps, precision=32)
    ctx = _record(sc, steps, every=8)
    p0 = _params(sc, proj)
    for dp in ([0.5, -0.25, 0.1, 3.0, 0.05], [-1.0, 0.3, 0.0, -5.0, -0.1]):
        p = p0 + np.array(dp)
        pos, rot = ctx.adjoint_replay(_options(sc, p[4]), _state_at(sc, proj, p))
        opos, orot = oracle.replay(sc.bodies, sc.state, rec, sc.options, p[4], precision=32, seed_body=proj,
                                   params=p)
        assert_bits(pos, opos, f"replay positions at {p}")
        assert_bits(rot, orot, f"replay orientations at {p}")
        _print("perturbed replay bit-exact at", p)
    ctx.close()


# ---------------------------------------------------------------- gradient
@pytest.mark.parametrize("steps,every", [(10, 4), (20, 8), (24, 5)])
def test_gradient_matches_float_valued_dual_numbers(oracle, steps, every):
    """Reverse mode (float32) against forward mode along the same float
    trajectory (float64 tangents): within 1e-3 of the largest component."""
    sc, proj = _throw()
    target = scenes.wall_throw_target(sc, 1)
    rec = oracle.record_world(sc, steps, precision=32)
    ctx = _record(sc, steps, every=every)
    p0 = _params(sc, proj)
    g_gpu, pos, _ = _gpu_grad(ctx, sc, proj, p0, target)
    g_ref, opos = _oracle_grad(oracle, sc, proj, rec, p0, target, 33)
    assert_bits(pos, opos, "replay positions")
    err = np.abs(g_gpu - g_ref)
    _print(f"steps {steps}: gpu {g_gpu} oracle {g_ref} |err| {err}")
    scale = np.max(np.abs(g_ref))
    assert scale > 0
    assert np.all(err <= 1e-3 * scale), (g_gpu, g_ref)
    ctx.close()


def test_reverse_launch_shapes_agree(oracle):
    """Every lanes-per-body shape of the reverse primal gives the same
    gradient (up to the order of the float sums over a body's rows)."""
    sc, proj = _throw()
    steps = 16
    target = scenes.wall_throw_target(sc, 1)
    rec = oracle.record_world(sc, steps, precision=32)
    ctx = _record(sc, steps, every=8)
    p0 = _params(sc, proj)
    g_ref, _ = _oracle_grad(oracle, sc, proj, rec, p0, target, 33)
    for lanes in (1, 2, 4, 8, 16):
        ctx.set_tuning(adjoint_lanes=lanes)
        g, _, _ = _gpu_grad(ctx, sc, proj, p0, target)
        _print("adjoint lanes", lanes, g)
        assert np.all(np.abs(g - g_ref) <= 1e-3 * np.max(np.abs(g_ref))), (lanes, g, g_ref)
    ctx.close()


def test_gradient_matches_float64_dual_numbers_and_differences(oracle):
    """Over 20 steps (the sphere has hit the wall) the float64 replay is
    smooth in v0 and m at the difference step 1e-4: the GPU gradient agrees
    with its dual numbers and central differences to 1e-3."""
    sc, proj = _throw()
    steps = 20
    target = scenes.wall_throw_target(sc, 1)
    rec = oracle.record_world(sc, steps, precision=32)
    ctx = _record(sc, steps)
    p0 = _params(sc, proj)
    g_gpu, _, _ = _gpu_grad(ctx, sc, proj, p0, target)
    g64, _ = _oracle_grad(oracle, sc, proj, rec, p0, target, 65)
    nb = sc.n - 1
    fd = np.zeros(4)
    for k in range(4):
        h = 1e-4 * max(abs(p0[k]), 1.0)
        e = np.eye(5)[k] * h
        lp = oracle.replay(sc.bodies, sc.state, rec, sc.options, 0.0, 64, proj, p0 + e)[0]
        lm = oracle.replay(sc.bodies, sc.state, rec, sc.options, 0.0, 64, proj, p0 - e)[0]
        lf = lambda x: float(np.mean(np.sum((x[:nb] - target) ** 2, axis=1)))
        fd[k] = (lf(lp) - lf(lm)) / (2 * h)
    _print("gpu", g_gpu[:4], "dual64", g64[:4], "fd", fd)
    scale = np.max(np.abs(g64[:4]))
    assert np.max(np.abs(g_gpu[:4] - g64[:4])) <= 1e-3 * scale
    assert np.max(np.abs(g_gpu[:4] - fd)) <= 1e-3 * scale
    ctx.close()


def test_full_wall_gradient_matches_float_valued_dual_numbers(oracle):
    """The 64 x 32 wall (2049 bodies) over 20 steps, seed 0: the reverse mode
    against forward-mode tangents along the same float trajectory.  The
    float64 replay's own derivative already differs by several per cent here
    (the Jacobi sweeps over the stiff wall amplify the difference between the
    float and double trajectories), so only the same-branch comparison is a
    check of the adjoint."""
    sc = scenes.wall_throw(0)
    proj = sc.n - 1
    steps = 20
    target = scenes.wall_throw_target(sc, 0)
    rec = oracle.record_world(sc, steps, precision=32)
    ctx = _record(sc, steps, every=8)
    p0 = _params(sc, proj)
    g_gpu, pos, _ = _gpu_grad(ctx, sc, proj, p0, target)
    g_ref, opos = _oracle_grad(oracle, sc, proj, rec, p0, target, 33)
    assert_bits(pos, opos, "replay positions")
    err = np.abs(g_gpu - g_ref)
    _print(f"64x32 wall, {steps} steps: gpu {g_gpu} oracle {g_ref} rel {err / np.abs(g_ref)}")
    assert np.all(err <= 1e-3 * np.abs(g_ref)), (g_gpu, g_ref)
    ctx.close()


def test_gradient_at_other_parameters(oracle):
    """Backward from state0: after a replay from the same state (checkpoints
    kept) and without one (a forward sweep first), both equal."""
    sc, proj = _throw()
    steps = 16
    target = scenes.wall_throw_target(sc, 1)
    rec = oracle.record_world(sc, steps, precision=32)
    ctx = _record(sc, steps, every=4)
    p = _params(sc, proj) + np.array([0.4, -0.2, 0.05, 2.0, 0.1])
    st = _state_at(sc, proj, p)
    g1, _, full1 = _gpu_grad(ctx, sc, proj, p, target, state0=st)
    # backward alone (the sweep key is dropped by a replay from another state)
    ctx.adjoint_replay(sc.options, None)
    nb = sc.n - 1
    opos, _ = oracle.replay(sc.bodies, sc.state, rec, sc.options, p[4], precision=32, seed_body=proj, params=p)
    dpos = np.zeros((sc.n, 3), np.float32)
    dpos[:nb] = (2.0 * (opos[:nb] - target) / nb).astype(np.float32)
    full2 = ctx.adjoint_backward(_options(sc, p[4]), d_positions=dpos, state0=st)
    for k in ("velocities", "inv_mass", "inv_moment", "positions"):
        assert_bits(full2[k], full1[k], f"backward {k} with and without the replay's checkpoints")
    assert full2["friction"] == full1["friction"]
    g_ref, _ = _oracle_grad(oracle, sc, proj, rec, p, target, 33)
    _print("at", p, "gpu", g1, "oracle", g_ref)
    assert np.all(np.abs(g1 - g_ref) <= 1e-3 * np.max(np.abs(g_ref)))
    ctx.close()


def test_batched_gradient_equals_per_scene():
    """Four throws batched in one world: each scene's adjoints equal those of
    the scene recorded alone, bit for bit (the friction adjoint, a sum over
    every contact, adds up the scenes)."""
    steps = 16
    throws = [scenes.wall_throw(s, columns=6, rows=4) for s in range(4)]
    batch = scenes.BatchedScenes(throws, np.zeros((0, 3), np.float32), throws[0].options)
    ctx = PhysicsContext(batch.n)
    ctx.batch_upload(batch)
    ctx.adjoint_begin(steps, 8)
    ctx.step(throws[0].options, steps)
    d = np.random.default_rng(0).normal(size=(batch.n, 3)).astype(np.float32)
    gb = ctx.adjoint_backward(throws[0].options, d_positions=d)
    mu_sum = 0.0
    for s, sc in enumerate(throws):
        lo, hi = batch.offsets[s], batch.offsets[s + 1]
        one = _record(sc, steps)
        g1 = one.adjoint_backward(sc.options, d_positions=d[lo:hi])
        for k in ("positions", "orientations", "velocities", "angular_velocities", "inv_mass", "inv_moment"):
            assert_bits(gb[k][lo:hi], g1[k], f"scene {s} {k}")
        mu_sum += g1["friction"]
        one.close()
    assert abs(gb["friction"] - mu_sum) <= 1e-5 * max(abs(mu_sum), 1e-12)
    ctx.close()


# ------------------------------------------------------ frozen multipliers
def _oracle_frozen_grad(oracle, sc, proj, rec, p, target, precision):
    """The oracle's frozen replay at p: the penalties and multipliers of the
    float run at p held fixed (what DIO_ADJOINT_FROZEN_MULTIPLIERS
    differentiates)."""
    nb = sc.n - 1
    pos, _, dpos, _ = oracle.replay_frozen(sc.bodies, sc.state, rec, sc.options, p[4], precision, proj, p, p, 32, 3)
    return np.einsum("ij,ijk->k", 2.0 * (pos[:nb] - target) / nb, dpos[:nb]), pos


@pytest.mark.parametrize("shape,seed,steps", [((6, 8), 1, 24), (None, 0, 40)])
def test_frozen_gradient_matches_float_valued_dual_numbers(oracle, shape, seed, steps):
    """DIO_ADJOINT_FROZEN_MULTIPLIERS: the reverse mode against the oracle's
    frozen replay with float values and float64 tangents (the same float
    trajectory and branches), within 1e-3 of the largest component, on the
    6 x 8 wall and on the 64 x 32 wall over 40 steps (where the exact
    derivative has already grown past 1e4).  Switching back to the exact mode
    on the same recording gives the exact derivative again."""
    sc = scenes.wall_throw(seed, columns=shape[0], rows=shape[1]) if shape else scenes.wall_throw(seed)
    proj = sc.n - 1
    target = scenes.wall_throw_target(sc, seed)
    rec = oracle.record_world(sc, steps, precision=32)
    ctx = _record(sc, steps, every=8)
    ctx.adjoint_set_mode("frozen_multipliers")
    p0 = _params(sc, proj)
    g_gpu, pos, _ = _gpu_grad(ctx, sc, proj, p0, target)
    g_ref, opos = _oracle_frozen_grad(oracle, sc, proj, rec, p0, target, 33)
    assert_bits(pos, opos, "replay positions")
    err = np.abs(g_gpu - g_ref)
    _print(f"frozen, {sc.n} bodies, {steps} steps: gpu {g_gpu} oracle {g_ref} |err| / max {err / np.max(np.abs(g_ref))}")
    assert np.all(np.isfinite(g_gpu)) and np.all(err <= 1e-3 * np.max(np.abs(g_ref))), (g_gpu, g_ref)
    if shape:
        ctx.adjoint_set_mode("exact")
        g_exact, _, _ = _gpu_grad(ctx, sc, proj, p0, target)
        g_ref_exact, _ = _oracle_grad(oracle, sc, proj, rec, p0, target, 33)
        assert np.all(np.abs(g_exact - g_ref_exact) <= 1e-3 * np.max(np.abs(g_ref_exact)))
        assert np.max(np.abs(g_exact - g_gpu)) > 1e-2 * np.max(np.abs(g_gpu))  # the modes differ here
    ctx.close()


def test_frozen_gradient_matches_float64_central_differences(oracle):
    """The 64 x 32 wall (2049 bodies), seed 0, over 20 steps: the GPU frozen
    gradient with respect to the projectile's v0 and mass against central
    differences at h = 1e-3 of the float64 frozen replay, each within 1e-3
    relative.  (The friction, and v0 past about 30 steps, are not smooth at
    h = 1e-3 in the float64 map itself: the CPU tests show the differences
    converging to the dual numbers only as h -> 0.)"""
    sc = scenes.wall_throw(0)
    proj = sc.n - 1
    nb = sc.n - 1
    steps = 20
    target = scenes.wall_throw_target(sc, 0)
    rec = oracle.record_world(sc, steps, precision=32)
    ctx = _record(sc, steps, every=8)
    ctx.adjoint_set_mode("frozen_multipliers")
    p0 = _params(sc, proj)
    g_gpu, _, _ = _gpu_grad(ctx, sc, proj, p0, target)
    lf = lambda x: float(np.mean(np.sum((x[:nb] - target) ** 2, axis=1)))
    fd = np.zeros(4)
    for k in range(4):
        e = np.eye(5)[k] * 1e-3
        lp = oracle.replay_frozen(sc.bodies, sc.state, rec, sc.options, 0.0, 64, proj, p0 + e, p0, 32, 3)[0]
        lm = oracle.replay_frozen(sc.bodies, sc.state, rec, sc.options, 0.0, 64, proj, p0 - e, p0, 32, 3)[0]
        fd[k] = (lf(lp) - lf(lm)) / 2e-3
    rel = np.abs(g_gpu[:4] - fd) / np.abs(fd)
    _print("frozen gpu", g_gpu[:4], "float64 differences (h = 1e-3)", fd, "rel", rel)
    assert np.all(rel <= 1e-3), rel
    ctx.close()


def test_frozen_gradient_over_120_steps_is_finite(oracle):
    """The configured throw length: 120 steps of the 64 x 32 wall.  The exact
    mode's float reverse mode overflows here; the frozen one stays finite and
    agrees with the oracle's frozen float-valued dual numbers."""
    sc = scenes.wall_throw(0)
    proj = sc.n - 1
    steps = 120
    target = scenes.wall_throw_target(sc, 0)
    rec = oracle.record_world(sc, steps, precision=32)
    ctx = _record(sc, steps, every=8)
    ctx.adjoint_set_mode("frozen_multipliers")
    p0 = _params(sc, proj)
    g_gpu, pos, _ = _gpu_grad(ctx, sc, proj, p0, target)
    g_ref, opos = _oracle_frozen_grad(oracle, sc, proj, rec, p0, target, 33)
    assert_bits(pos, opos, "replay positions")
    err = np.abs(g_gpu - g_ref)
    _print(f"frozen, 120 steps: gpu {g_gpu} oracle {g_ref} |err| / max {err / np.max(np.abs(g_ref))}")
    assert np.all(np.isfinite(g_gpu))
    assert np.all(err <= 1e-2 * np.max(np.abs(g_ref))), (g_gpu, g_ref)
    ctx.close()


# ------------------------------------------------------------ torch binding
def test_torch_autograd_binding(oracle):
    import torch

    from diomedes_amd.autograd import replay

    sc, proj = _throw()
    steps = 12
    target = torch.from_numpy(scenes.wall_throw_target(sc, 1))
    ctx = _record(sc, steps, every=4)
    vel = torch.from_numpy(sc.state["vel"][:, :3].copy()).requires_grad_()
    ang = torch.from_numpy(sc.state["ang"][:, :3].copy()).requires_grad_()
    im = torch.from_numpy(sc.state["inv_mass"].copy()).requires_grad_()
    imo = torch.from_numpy(sc.state["inv_moment"][:, :3].copy()).requires_grad_()
    mu = torch.tensor(float(sc.options.friction), requires_grad=True)
    pos, rot = replay(ctx, sc.options, sc.state, vel, ang, im, imo, mu)
    nb = sc.n - 1
    loss = ((pos[:nb].double() - target) ** 2).sum(dim=1).mean()
    loss.backward()
    dpos = np.zeros((sc.n, 3), np.float32)
    dpos[:nb] = (2.0 * (pos[:nb].detach().double().numpy() - target.numpy()) / nb).astype(np.float32)
    g = ctx.adjoint_backward(sc.options, d_positions=dpos, state0=sc.state)
    assert_bits(vel.grad.numpy(), g["velocities"], "torch velocity grad")
    assert_bits(im.grad.numpy(), g["inv_mass"], "torch inverse-mass grad")
    assert abs(float(mu.grad) - np.float32(g["friction"])) == 0.0
    # friction as a plain number takes no gradient; as a shape-(1,) tensor
    # its gradient has that shape
    vel2 = torch.from_numpy(sc.state["vel"][:, :3].copy()).requires_grad_()
    pos2, _ = replay(ctx, sc.options, sc.state, vel2, ang.detach(), im.detach(), imo.detach(), float(sc.options.friction))
    ((pos2[:nb].double() - target) ** 2).sum(dim=1).mean().backward()
    assert_bits(vel2.grad.numpy(), g["velocities"], "torch velocity grad, friction a number")
    mu1 = torch.tensor([float(sc.options.friction)], requires_grad=True)
    pos3, _ = replay(ctx, sc.options, sc.state, vel.detach(), ang.detach(), im.detach(), imo.detach(), mu1)
    ((pos3[:nb].double() - target) ** 2).sum(dim=1).mean().backward()
    assert mu1.grad.shape == (1,) and float(mu1.grad[0]) == float(mu.grad)
    ctx.close()


# ------------------------------------------------------------- error paths
def test_adjoint_errors():
    from diomedes_amd._lib import DioError

    sc, _ = _throw(rows=2, columns=2)
    with PhysicsContext(sc.n) as ctx:
        ctx.upload(sc)
        with pytest.raises(DioError):
            ctx.adjoint_replay(sc.options)  # nothing recorded
        ctx.adjoint_begin(3, 2)
        ctx.step(sc.options, 3)
        assert ctx.adjoint_steps() == 3
        ctx.step(sc.options, 1)  # beyond max_steps: flagged, replay refuses
        with pytest.raises(DioError):
            ctx.adjoint_replay(sc.options)
        ctx.adjoint_begin(4, 2)
        with pytest.raises(DioError):
            ctx.adjoint_backward(sc.options)  # no steps recorded yet
        ctx.step(sc.options, 1)
        g = ctx.adjoint_backward(sc.options)  # a zero final adjoint: all zero
        assert all(np.all(g[k] == 0.0) for k in ("positions", "velocities", "inv_mass")) and g["friction"] == 0.0
        with pytest.raises(ValueError):
            ctx.adjoint_set_mode("smoothed")
        with pytest.raises(DioError):
            ctx._check(ctx._lib.dio_adjoint_set_mode(ctx._h, 2))
        ctx.upload(sc)  # a new world frees the recording
        with pytest.raises(DioError):
            ctx.adjoint_steps()

"""GPU coloured Gauss-Seidel mode (dio_step mode 1) against the oracle's.

The colouring is integer work and must equal the oracle's exactly; the
coloured step, like the Jacobi step, must be bit-identical to the FP32
oracle's coloured step (same primal arithmetic, bodies of one colour share no
contact or joint, so the order within a colour does not matter).
"""

from __future__ import annotations

import numpy as np
import pytest

from diomedes_amd import PhysicsContext, scenes
from diomedes_amd.physics import AvbdBody, AvbdOptions, Scene, Spring

from .helpers import assert_bits

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _compare(ctx, w, what):
    g = ctx.download()
    o = w.state()
    for key in ("positions", "orientations", "velocities", "angular_velocities"):
        assert_bits(g[key], o[key], f"{what}: {key}")
    gc, gn = ctx.colours()
    oc, on, _ = w.colours()
    assert gn == on, (what, gn, on)
    assert np.array_equal(gc, oc), what


def _scene(name):
    if name == "stack":
        return scenes.box_stack(0, 10)
    if name == "field":
        return scenes.cube_field(1, nx=8, ny=3, nz=8)
    if name == "mixed":
        return scenes.mixed_field(1, n=400, ropes=4, rope_links=8, springs=16)
    if name == "wall":
        return scenes.wall_throw(2, columns=16, rows=8)
    raise ValueError(name)


@pytest.mark.parametrize("name,lanes", [("stack", 32), ("field", 32), ("field", 1), ("field", 8), ("mixed", 32),
                                        ("wall", 16)])
def test_coloured_step_bitexact(oracle, name, lanes):
    sc = _scene(name)
    w = oracle.OracleWorld(sc, 32)
    w.set_mode(1)
    with PhysicsContext(sc.n, max_hull_vertices=len(sc.hull)) as ctx:
        ctx.set_tuning(primal_lanes=lanes)
        ctx.upload(sc)
        for k in range(3):
            ctx.step(sc.options, 15, mode=1)
            w.step(nsteps=15)
            _compare(ctx, w, f"{name} step {15 * (k + 1)}")
            st = ctx.stats()
            _print(name, "step", 15 * (k + 1), "colours", st.colours, "rounds", st.colour_rounds, "contacts",
                   st.contacts)


@pytest.mark.parametrize("lanes", [32, 1])
def test_captured_coloured_step_equals_uncaptured(lanes, monkeypatch):
    """The coloured segment is captured after its first step, with caps on the
    colouring rounds and on the colours launched one by one taken from that
    step; later steps that need more (the field starts falling apart, with one
    colour, and piles up) finish the colouring in one workgroup and run the
    extra colours one by one.  Every step equals the uncaptured path (no
    graphs) bit for bit, and the colour and round counts agree."""
    sc = _scene("field")
    res = {}
    for graphs in (True, False):
        if graphs:
            monkeypatch.delenv("DIO_NO_GRAPH", raising=False)
        else:
            monkeypatch.setenv("DIO_NO_GRAPH", "1")
        with PhysicsContext(sc.n) as ctx:
            ctx.set_tuning(primal_lanes=lanes)
            ctx.upload(sc)
            out = []
            for chunk in (1, 10, 25, 25):
                ctx.step(sc.options, chunk, mode=1)
                st = ctx.stats()
                out.append((ctx.download(), ctx.colours(), st.colours, st.colour_rounds))
            res[graphs] = out
    for k, ((g, gc, gn, gr), (u, uc, un, ur)) in enumerate(zip(res[True], res[False])):
        for key in ("positions", "orientations", "velocities", "angular_velocities"):
            assert_bits(g[key], u[key], f"call {k} {key}")
        assert np.array_equal(gc[0], uc[0]) and gc[1] == uc[1]
        assert (gn, gr) == (un, ur), (k, gn, gr, un, ur)
        _print(f"lanes {lanes} call {k}: colours {gn}, rounds {gr}: captured = uncaptured")
    assert res[True][0][2] == 1 and res[True][-1][2] > 2  # the colour count grew past the first capture's cap


def test_colour_lanes_apart_from_primal_lanes(oracle):
    """colour_lanes sets the coloured primal's shape on its own (here 4 lanes
    per body against the Jacobi primal's 32): still the oracle's step."""
    sc = _scene("field")
    w = oracle.OracleWorld(sc, 32)
    with PhysicsContext(sc.n) as ctx:
        ctx.set_tuning(primal_lanes=32, colour_lanes=4)
        ctx.upload(sc)
        for mode in (0, 1, 1):
            w.set_mode(mode)
            ctx.step(sc.options, 12, mode=mode)
            w.step(nsteps=12)
        _compare(ctx, w, "colour lanes 4")
        assert ctx.tuning()["colour_lanes"] == 4


def test_modes_alternate(oracle):
    """Jacobi and coloured steps can follow each other on one world."""
    sc = _scene("field")
    w = oracle.OracleWorld(sc, 32)
    with PhysicsContext(sc.n) as ctx:
        ctx.upload(sc)
        for k, mode in enumerate((0, 1, 0, 1)):
            w.set_mode(mode)
            ctx.step(sc.options, 8, mode=mode)
            w.step(nsteps=8)
            g, o = ctx.download(), w.state()
            assert_bits(g["positions"], o["positions"], f"phase {k} mode {mode}")


def test_coloured_chain_converges_as_the_oracle(oracle):
    bodies = [AvbdBody.sphere((0.0, 10.0, 0.0), 0.1, 0.0)]
    for i in range(12):
        bodies.append(AvbdBody.sphere((0.0, 10.0 - 0.65 * (i + 1), 0.0), 0.1, 1.0))
    cons = [Spring(i, i + 1, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 1.0e4, 0.5) for i in range(12)]
    o = AvbdOptions(iterations=8, gravity=(0.0, -10.0, 0.0))
    sc = Scene.from_bodies(bodies, cons, None, o)
    w = oracle.OracleWorld(sc, 32)
    w.set_mode(1)
    with PhysicsContext(sc.n, max_constraints=len(cons)) as ctx:
        ctx.upload(sc)
        ctx.step(o, 5, mode=1)
        w.step(o, nsteps=5)
        _compare(ctx, w, "chain")
        assert ctx.stats().colours == 2


def test_coloured_batched_matches_plain():
    batch = scenes.batched_scenes(3, 24, n_min=8, n_max=96)
    with PhysicsContext(batch.n, max_pairs=8 * batch.n, max_contacts=16 * batch.n,
                        max_hull_vertices=len(batch.hull)) as ctx:
        ctx.batch_upload(batch)
        ctx.step(batch.options, 20, mode=1)
        gb = ctx.download()
    for s in (0, 7, 23):
        sc = batch.scenes[s]
        with PhysicsContext(sc.n, max_hull_vertices=len(sc.hull)) as one:
            one.upload(sc)
            one.step(batch.options, 20, mode=1)
            g1 = one.download()
        sl = batch.scene_slice(s)
        assert_bits(gb["positions"][sl], g1["positions"], f"scene {s}")


def test_coloured_mode_errors():
    from diomedes_amd._lib import DioError

    sc = _scene("stack")
    with PhysicsContext(sc.n) as ctx:
        ctx.upload(sc)
        with pytest.raises(DioError):
            ctx.step(sc.options, 1, mode=2)
        ctx.adjoint_begin(4, 2)
        with pytest.raises(DioError):
            ctx.step(sc.options, 1, mode=1)
        ctx.adjoint_end()
        ctx.step(sc.options, 2, mode=1)

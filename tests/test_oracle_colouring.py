"""The oracle's graph colouring and coloured Gauss-Seidel solve (no
counterpart in the reference, whose solve is Jacobi).

The colouring is pinned by what any Jones-Plassmann greedy colouring must
satisfy: no edge inside a colour, every vertex of colour c has neighbours of
every colour below c, at most max-degree + 1 colours, independence of edge
order and duplicates, and the closed cases (no edges, a complete graph, a
star, a path).  The Gauss-Seidel solve is pinned by the theory of the two
iterations on a spring chain (a two-colourable, consistently ordered system):
both converge to the same implicit step, and Gauss-Seidel's error after I
iterations matches Jacobi's after 2I (rho_GS = rho_J^2).
"""

from __future__ import annotations

import numpy as np
import pytest

from diomedes_amd import scenes
from diomedes_amd.physics import AvbdBody, AvbdOptions, Scene, Spring

NOT_IN_GRAPH = 0xFFFFFFFE


def _check_colouring(n, dyn, edges, col, nc):
    dyn = np.asarray(dyn, bool)
    assert np.all(col[~dyn] == NOT_IN_GRAPH)
    assert np.all(col[dyn] < 255)
    adj = [set() for _ in range(n)]
    for a, b in edges:
        if a != b and dyn[a] and dyn[b]:
            adj[a].add(b)
            adj[b].add(a)
    for a in range(n):
        if not dyn[a]:
            continue
        for b in adj[a]:
            assert col[a] != col[b], (a, b)
        below = {int(col[b]) for b in adj[a]}
        assert set(range(int(col[a]))) <= below, (a, col[a], below)
    maxdeg = max((len(s) for s in adj), default=0)
    used = col[dyn]
    assert nc == (int(used.max()) + 1 if used.size else 0)
    assert nc <= maxdeg + 1


@pytest.mark.parametrize("seed,n,deg,static_frac", [(0, 300, 6, 0.1), (1, 2000, 12, 0.0), (2, 50, 30, 0.3)])
def test_random_graph_colouring(oracle, seed, n, deg, static_frac):
    rng = np.random.default_rng(seed)
    edges = rng.integers(0, n, size=(n * deg // 2, 2))
    dyn = rng.random(n) >= static_frac
    col, nc, rounds = oracle.colour_graph(n, dyn, edges)
    _check_colouring(n, dyn, edges, col, nc)
    # the same colouring for shuffled, reversed and duplicated edges
    e2 = np.concatenate([edges[rng.permutation(len(edges))][:, ::-1], edges[: len(edges) // 3]])
    col2, nc2, rounds2 = oracle.colour_graph(n, dyn, e2)
    assert np.array_equal(col, col2) and nc == nc2 and rounds == rounds2


def test_closed_cases(oracle):
    col, nc, rounds = oracle.colour_graph(0, [], np.zeros((0, 2)))
    assert nc == 0 and rounds == 0
    col, nc, rounds = oracle.colour_graph(7, np.ones(7), np.zeros((0, 2)))
    assert nc == 1 and rounds == 1 and np.all(col == 0)
    k = 12
    complete = np.array([(a, b) for a in range(k) for b in range(a + 1, k)])
    col, nc, _ = oracle.colour_graph(k, np.ones(k), complete)
    assert nc == k and sorted(col.tolist()) == list(range(k))
    star = np.array([(0, b) for b in range(1, 40)])
    col, nc, _ = oracle.colour_graph(40, np.ones(40), star)
    assert nc == 2 and len(set(col[1:].tolist())) == 1 and col[0] != col[1]
    path = np.array([(a, a + 1) for a in range(99)])
    col, nc, _ = oracle.colour_graph(100, np.ones(100), path)
    _check_colouring(100, np.ones(100), path, col, nc)
    assert nc <= 3
    col, nc, rounds = oracle.colour_graph(5, np.zeros(5), star[:4])
    assert nc == 0 and np.all(col == NOT_IN_GRAPH)


def _contact_edges(w):
    c = w.contacts()
    ab = np.stack([c["a"], c["b"]], axis=1).astype(np.int64)
    return ab[ab[:, 1] != 0xFFFFFFFF]


@pytest.mark.parametrize("name", ["stack", "field", "mixed"])
def test_world_colouring_separates_coupled_bodies(oracle, name):
    if name == "stack":
        sc = scenes.box_stack(0, 10)
    elif name == "field":
        sc = scenes.cube_field(0, nx=6, ny=3, nz=6)
    else:
        sc = scenes.mixed_field(1, n=300, ropes=4, rope_links=8, springs=16)
    w = oracle.OracleWorld(sc, 32)
    w.set_mode(1)
    for _ in range(3):
        w.step(nsteps=20)
        col, nc, rounds = w.colours()
        dyn = sc.state["inv_mass"] != 0
        edges = _contact_edges(w)
        if len(sc.constraints):
            edges = np.concatenate([edges, np.stack([sc.constraints["a"], sc.constraints["b"]], 1).astype(np.int64)])
        _check_colouring(sc.n, dyn, edges, col, nc)
        assert nc >= 2 and rounds >= 1


def _chain(k, nb=12, stretch=1.3):
    bodies = [AvbdBody.sphere((0.0, 10.0, 0.0), 0.1, 0.0)]
    for i in range(nb):
        bodies.append(AvbdBody.sphere((0.0, 10.0 - 0.5 * stretch * (i + 1), 0.0), 0.1, 1.0))
    cons = [Spring(i, i + 1, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), k, 0.5) for i in range(nb)]
    return bodies, cons


def _chain_step(oracle, k, iters, mode):
    bodies, cons = _chain(k)
    o = AvbdOptions(iterations=iters, gravity=(0.0, -10.0, 0.0))
    sc = Scene.from_bodies(bodies, cons, None, o)
    w = oracle.OracleWorld(sc, 64)
    w.set_mode(mode)
    w.step(o, nsteps=1)
    return w.state()["positions"][:, 1]


def test_gauss_seidel_converges_twice_as_fast_on_a_chain(oracle):
    k = 1.0e4
    y_gs = _chain_step(oracle, k, 1000, 1)
    y_j = _chain_step(oracle, k, 1000, 0)
    assert np.max(np.abs(y_gs - y_j)) < 1e-12  # the same implicit step
    for iters in (2, 4, 8, 16):
        e_gs = np.max(np.abs(_chain_step(oracle, k, iters, 1) - y_gs))
        e_j = np.max(np.abs(_chain_step(oracle, k, iters, 0) - y_gs))
        e_j2 = np.max(np.abs(_chain_step(oracle, k, 2 * iters, 0) - y_gs))
        assert e_gs < e_j
        assert 0.5 < e_gs / e_j2 < 2.0, (iters, e_gs, e_j2)


def test_gauss_seidel_without_coupling_is_jacobi(oracle):
    """Free-falling bodies share no contact: one colour, and the coloured
    solve is the Jacobi solve bit for bit."""
    bodies = [AvbdBody.cube((3.0 * i, 5.0 + i, 0.0), 0.5, 1.0) for i in range(6)]
    for i, b in enumerate(bodies):
        b.velocity = (0.3 * i, 1.0, -0.2)
        b.angular_velocity = (0.1, 0.2 * i, 0.3)
    sc = Scene.from_bodies(bodies, (), None, AvbdOptions(iterations=4))
    out = []
    for mode in (0, 1):
        w = oracle.OracleWorld(sc, 32)
        w.set_mode(mode)
        w.step(nsteps=30)
        out.append(w.state())
    for key in ("positions", "orientations", "velocities", "angular_velocities"):
        assert np.array_equal(out[0][key], out[1][key]), key

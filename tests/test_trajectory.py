"""Scene and trajectory files (diomedes_amd.trajectory): byte layout, round
trips and error handling on the CPU; the oracle's FP32 trajectory of the
10-cube stack against the committed golden file."""

from __future__ import annotations

import struct
from pathlib import Path

import numpy as np
import pytest

from diomedes_amd import scenes, trajectory as T

GOLDEN = Path(__file__).resolve().parent / "golden" / "box_stack_s0.dtraj"


def _frames(n, count, rng, flags):
    out = []
    for k in range(count):
        f = T.Frame(10 * k, rng.normal(size=(n, 3)).astype(np.float32), rng.normal(size=(n, 4)).astype(np.float32))
        if flags & T.FLAG_VELOCITIES:
            f.velocities = rng.normal(size=(n, 3)).astype(np.float32)
            f.angular_velocities = rng.normal(size=(n, 3)).astype(np.float32)
        if flags & T.FLAG_PAIRS:
            f.pairs = np.sort(rng.integers(0, n, size=(k * 3, 2)).astype(np.uint32), axis=0)
        out.append(f)
    return out


def _same_scene(a, b):
    for k in ("bodies", "state", "constraints"):
        assert getattr(a, k).tobytes() == getattr(b, k).tobytes(), k
    assert np.array_equal(a.hull, b.hull)
    assert bytes(a.options.to_c()) == bytes(b.options.to_c())


@pytest.mark.parametrize("flags", [0, T.FLAG_VELOCITIES, T.FLAG_PAIRS, T.FLAG_VELOCITIES | T.FLAG_PAIRS])
def test_round_trip(tmp_path, flags):
    sc = scenes.mixed_field(2, n=60, ropes=2, rope_links=4, springs=6)
    rng = np.random.default_rng(flags)
    tr = T.Trajectory(sc, np.array([0, sc.n], np.uint32), [sc.container], flags, _frames(sc.n, 4, rng, flags))
    p = tmp_path / "t.dtraj"
    T.write(p, tr)
    back = T.read(p)
    _same_scene(sc, back.scene)
    assert back.flags == flags and len(back.frames) == 4
    c0, c1 = sc.container, back.containers[0]
    assert tuple(np.float32(c0.center)) == c1.center and np.float32(c0.half_extent) == c1.half_extent
    for f, g in zip(tr.frames, back.frames):
        assert f.step == g.step
        for k in ("positions", "orientations", "velocities", "angular_velocities", "pairs"):
            a, b = getattr(f, k), getattr(g, k)
            assert (a is None and b is None) or np.array_equal(a, b), k
    # the byte layout: header, then the options record, then the bodies
    raw = p.read_bytes()
    assert raw[:8] == b"DIOTRAJ\x00"
    assert struct.unpack_from("<6I", raw, 8) == (1, sc.n, len(sc.constraints), len(sc.hull), 1, flags)
    assert struct.unpack_from("<5I", raw, 32) == (112, 80, 48, 48, 36)
    assert raw[64:112] == bytes(sc.options.to_c())
    assert raw[112:112 + 112 * sc.n] == sc.bodies.tobytes()


def test_batched_scene_table(tmp_path):
    batch = scenes.batched_scenes(1, 5, n_min=8, n_max=20)
    from diomedes_amd.physics import Scene

    sc = Scene(batch.bodies, batch.state, hull=batch.hull, options=batch.options)
    p = tmp_path / "b.dtraj"
    with T.Writer(p, sc, 0, batch.offsets, batch.containers) as w:
        w.frame(0, batch.bodies["position"][:, :3], batch.bodies["orientation"])
    back = T.read(p)
    assert np.array_equal(back.scene_offsets, batch.offsets)
    assert len(back.containers) == 5 and all(c is not None for c in back.containers)
    assert back.scene.container is None
    _same_scene(sc, back.scene)


def test_rejects_bad_files(tmp_path):
    sc = scenes.box_stack(0)
    p = tmp_path / "ok.dtraj"
    with T.Writer(p, sc) as w:
        w.frame(1, np.zeros((sc.n, 3)), np.zeros((sc.n, 4)))
    raw = bytearray(p.read_bytes())
    for name, mutate, msg in (
        ("magic", lambda b: b.__setitem__(slice(0, 8), b"NOTTRAJ\x00"), "magic"),
        ("version", lambda b: b.__setitem__(slice(8, 12), struct.pack("<I", 9)), "version"),
        ("sizes", lambda b: b.__setitem__(slice(32, 36), struct.pack("<I", 100)), "record sizes"),
    ):
        bad = bytearray(raw)
        mutate(bad)
        q = tmp_path / f"{name}.dtraj"
        q.write_bytes(bytes(bad))
        with pytest.raises(ValueError, match=msg):
            T.read(q)
    q = tmp_path / "short.dtraj"
    q.write_bytes(bytes(raw[:-5]))
    with pytest.raises(ValueError, match="truncated"):
        T.read(q)
    with pytest.raises(ValueError):
        with T.Writer(tmp_path / "x.dtraj", sc, T.FLAG_PAIRS) as w:
            w.frame(0, np.zeros((sc.n, 3)), np.zeros((sc.n, 4)))


def test_oracle_reproduces_the_golden_trajectory(oracle):
    g = T.read(GOLDEN)
    assert g.flags == T.FLAG_VELOCITIES | T.FLAG_PAIRS
    assert [f.step for f in g.frames] == list(range(0, 301, 10))
    sc = scenes.box_stack(0)
    _same_scene(sc, g.scene)
    w = oracle.OracleWorld(sc, 32)
    done = 0
    for f in g.frames[1:]:
        w.step(nsteps=f.step - done)
        done = f.step
        st = w.state()
        assert np.array_equal(st["positions"].astype(np.float32), f.positions)
        assert np.array_equal(st["orientations"].astype(np.float32), f.orientations)
        assert np.array_equal(st["velocities"].astype(np.float32), f.velocities)
        assert np.array_equal(w.pairs(), f.pairs)


@pytest.mark.gpu
def test_gpu_records_the_golden_trajectory(tmp_path):
    """The fused device step written to a trajectory file by
    trajectory.record equals the golden file of the FP32 oracle, bit for bit
    (poses, velocities and pairs every 10 steps over 300)."""
    from diomedes_amd import PhysicsContext

    sc = scenes.box_stack(0)
    p = tmp_path / "gpu.dtraj"
    with PhysicsContext(sc.n) as ctx:
        frames = T.record(ctx, sc, p, 300, every=10, velocities=True, pairs=True)
    assert frames == 31
    got, ref = T.read(p), T.read(GOLDEN)
    _same_scene(ref.scene, got.scene)
    for f, g in zip(ref.frames, got.frames):
        assert f.step == g.step
        for k in ("positions", "orientations", "velocities", "angular_velocities", "pairs"):
            assert np.array_equal(getattr(f, k), getattr(g, k)), (f.step, k)



@pytest.mark.gpu
def test_gpu_record_names_the_options_it_steps_with(tmp_path, oracle):
    """trajectory.record with options other than the scene's writes those
    options into the header, so the file replays: the oracle stepping the
    file's own scene and options reproduces its frames."""
    import dataclasses

    from diomedes_amd import PhysicsContext

    sc = scenes.box_stack(0)
    opts = dataclasses.replace(sc.options, iterations=6, friction=0.3)
    p = tmp_path / "opts.dtraj"
    with PhysicsContext(sc.n) as ctx:
        T.record(ctx, sc, p, 40, every=20, options=opts)
    got = T.read(p)
    assert bytes(got.scene.options.to_c()) == bytes(opts.to_c())
    w = oracle.OracleWorld(got.scene, 32)
    for f, k in zip(got.frames[1:], (20, 20)):
        w.step(got.scene.options, nsteps=k)
        assert np.array_equal(w.state()["positions"].astype(np.float32), f.positions)

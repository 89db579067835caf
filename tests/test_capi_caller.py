"""A plain C99 program (tests/capi/replay_trajectory.c) drives the world path
through the C ABI with no Python or C++ in the caller: it reads a scene and
trajectory file, uploads the scene, steps it to every frame and compares
poses, velocities and broad-phase pairs bit for bit, then checks the status
contract (DIO_INVALID, DIO_CAPACITY with the required size, DIO_NOT_READY).

The trajectories it replays are the FP32 oracle's: the committed golden file
of the 10-cube stack, and files the oracle writes here for a small convex
field with hulls and joints, a batched table of scenes (each scene stepped
alone by the oracle, frames concatenated) and hard-joint chains in the
coloured mode.  The CPU tests run the program's file reader only.
"""

from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from diomedes_amd import records as R, scenes, trajectory as T
from diomedes_amd.build import CAPI_DIR, CAPI_EXE, build_capi_caller
from diomedes_amd.physics import Scene

GOLDEN = Path(__file__).resolve().parent / "golden" / "box_stack_s0.dtraj"


@pytest.fixture(scope="module")
def exe():
    src = CAPI_DIR / "replay_trajectory.c"
    if not CAPI_EXE.exists() or CAPI_EXE.stat().st_mtime < src.stat().st_mtime:
        build_capi_caller(verbose=False)
    return CAPI_EXE


def _run(exe, *args, timeout=120):
    p = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=timeout)
    print(p.stdout + p.stderr, flush=True)
    return p


def _mixed_small():
    return scenes.mixed_field(3, n=120, ropes=2, rope_links=5, springs=8, extent=6.0)


def _oracle_file(oracle, path, scene, steps, every, mode=0):
    """The FP32 oracle's trajectory of one scene, frames every `every` steps
    with velocities and pairs."""
    w = oracle.OracleWorld(scene, 32)
    w.set_mode(mode)
    with T.Writer(path, scene, T.FLAG_VELOCITIES | T.FLAG_PAIRS) as f:
        done = 0
        while True:
            st = w.state()
            f.frame(done, st["positions"], st["orientations"], st["velocities"], st["angular_velocities"],
                    w.pairs() if done else np.zeros((0, 2), np.uint32))
            if done >= steps:
                break
            w.step(nsteps=every)
            done += every


def _oracle_batch_file(oracle, path, batch, steps, every):
    """A batched table: each scene stepped alone by the oracle, each frame the
    concatenation (pairs in global indices, still sorted: scenes hold
    increasing index ranges)."""
    worlds = [oracle.OracleWorld(sc, 32) for sc in batch.scenes]
    whole = Scene(batch.bodies, batch.state, hull=batch.hull, options=batch.options)
    with T.Writer(path, whole, T.FLAG_VELOCITIES | T.FLAG_PAIRS, batch.offsets, batch.containers) as f:
        done = 0
        while True:
            sts = [w.state() for w in worlds]
            cat = {k: np.concatenate([s[k] for s in sts]) for k in sts[0]}
            if done:
                pairs = np.concatenate([np.asarray(w.pairs(), np.uint32).reshape(-1, 2) + np.uint32(o)
                                        for w, o in zip(worlds, batch.offsets[:-1])])
            else:
                pairs = np.zeros((0, 2), np.uint32)
            f.frame(done, cat["positions"], cat["orientations"], cat["velocities"], cat["angular_velocities"], pairs)
            if done >= steps:
                break
            for w in worlds:
                w.step(nsteps=every)
            done += every


def test_caller_reads_the_golden_file(exe):
    p = _run(exe, "--parse-only", GOLDEN)
    assert p.returncode == 0
    assert "10 bodies, 0 constraints, 0 hull vertices, 1 scenes, flags 3" in p.stdout
    assert "31 frames, last step 300" in p.stdout


def test_caller_reads_hulls_joints_and_scene_tables(exe, tmp_path):
    sc = _mixed_small()
    a = tmp_path / "mixed.dtraj"
    with T.Writer(a, sc, T.FLAG_VELOCITIES) as w:
        for k in range(3):
            w.frame(5 * k, sc.bodies["position"][:, :3], sc.bodies["orientation"], sc.state["vel"][:, :3],
                    sc.state["ang"][:, :3])
    p = _run(exe, "--parse-only", a)
    assert p.returncode == 0
    assert f"{sc.n} bodies, {len(sc.constraints)} constraints, {len(sc.hull)} hull vertices, 1 scenes, flags 1" \
        in p.stdout
    assert "3 frames, last step 10" in p.stdout
    batch = scenes.batched_scenes(1, 5, n_min=8, n_max=20)
    b = tmp_path / "batch.dtraj"
    with T.Writer(b, Scene(batch.bodies, batch.state, hull=batch.hull, options=batch.options), T.FLAG_PAIRS,
                  batch.offsets, batch.containers) as w:
        w.frame(0, batch.bodies["position"][:, :3], batch.bodies["orientation"], pairs=np.zeros((0, 2), np.uint32))
        w.frame(7, batch.bodies["position"][:, :3], batch.bodies["orientation"], pairs=np.array([[0, 1], [2, 5]]))
    p = _run(exe, "--parse-only", b)
    assert p.returncode == 0
    assert f"{batch.n} bodies, 0 constraints, {len(batch.hull)} hull vertices, 5 scenes, flags 2" in p.stdout
    assert "2 frames, last step 7" in p.stdout


def test_caller_rejects_bad_files(exe, tmp_path):
    raw = GOLDEN.read_bytes()
    cases = {
        "magic": b"NOTTRAJ\x00" + raw[8:],
        "version": raw[:8] + (9).to_bytes(4, "little") + raw[12:],
        "sizes": raw[:32] + (100).to_bytes(4, "little") + raw[36:],
        "truncated_scene": raw[:200],
        "truncated_frame": raw[:-5],
        "frame_tag": raw[:64 + 48 + 10 * (112 + 80) + 8 + 36] + b"XXXX" + raw[64 + 48 + 10 * (112 + 80) + 8 + 36 + 4:],
    }
    for name, data in cases.items():
        q = tmp_path / f"{name}.dtraj"
        q.write_bytes(data)
        p = _run(exe, "--parse-only", q)
        assert p.returncode == 3, name
        assert "bad trajectory file" in p.stderr, name
    assert _run(exe).returncode == 3
    assert _run(exe, "--parse-only", tmp_path / "missing.dtraj").returncode == 3


@pytest.mark.gpu
def test_c_caller_replays_the_golden_trajectory(exe):
    p = _run(exe, GOLDEN)
    assert p.returncode == 0, p.stderr
    assert "replayed 31 frames to step 300: poses, velocities, pairs bit-identical" in p.stdout
    assert "DIO_INVALID" in p.stdout and "DIO_CAPACITY" in p.stdout and "DIO_NOT_READY" in p.stdout


@pytest.mark.gpu
def test_c_caller_replays_a_convex_field_with_joints(exe, oracle, tmp_path):
    sc = _mixed_small()
    assert len(sc.hull) > 0 and len(sc.constraints) > 0
    f = tmp_path / "mixed.dtraj"
    _oracle_file(oracle, f, sc, 200, 25)
    p = _run(exe, f)
    assert p.returncode == 0, p.stderr
    assert "replayed 9 frames to step 200" in p.stdout


@pytest.mark.gpu
def test_c_caller_replays_a_batched_table(exe, oracle, tmp_path):
    batch = scenes.batched_scenes(2, 6, n_min=8, n_max=40)
    f = tmp_path / "batch.dtraj"
    _oracle_batch_file(oracle, f, batch, 40, 8)
    p = _run(exe, f)
    assert p.returncode == 0, p.stderr
    assert "replayed 6 frames to step 40" in p.stdout


@pytest.mark.gpu
def test_c_caller_replays_hard_chains_coloured(exe, oracle, tmp_path):
    sc = scenes.hard_chains(5, chains=6, links=6)
    assert np.any(sc.constraints["kind"] == R.HARD_SOCKET) and np.any(sc.constraints["kind"] == R.HARD_ROD)
    f = tmp_path / "chains.dtraj"
    _oracle_file(oracle, f, sc, 45, 15, mode=1)
    p = _run(exe, "--mode", 1, f)
    assert p.returncode == 0, p.stderr
    assert "replayed 4 frames to step 45" in p.stdout

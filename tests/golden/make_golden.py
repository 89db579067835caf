"""Write the golden fixtures of the FP32 / FP64 oracle.

python tests/golden/make_golden.py

- box_stack_s0.npz: the 10-cube stack (seed 0), 300 steps of the step
  protocol (contact lag 1); poses and velocities every 10 steps, and the last
  step's contacts and pairs, from the FP32 and the FP64 oracle.
- cube_field_8x2x8_s7.npz: 128 random cubes (seed 7), 120 steps, poses and
  velocities every 20 steps (FP32 and FP64).
- box_stack_s0.dtraj: the 10-cube stack's FP32 trajectory as a scene and
  trajectory file: frames every 10 steps with velocities and pairs.

The fixtures pin the oracle build against regressions and let GPU tests check
trajectories without recomputing them.  They are outputs of this build's
oracle (no reference run is possible: the reference is Rust / rust-gpu and
needs a toolchain this environment lacks).
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from diomedes_amd import scenes  # noqa: E402
from oracle import pyoracle  # noqa: E402


def trajectory(scene, steps: int, every: int, precision: int) -> dict:
    w = pyoracle.OracleWorld(scene, precision)
    keys = ("positions", "orientations", "velocities", "angular_velocities")
    out = {k: [] for k in keys}
    for _ in range(steps // every):
        w.step(nsteps=every)
        st = w.state()
        for k in keys:
            out[k].append(st[k])
    res = {k: np.array(v) for k, v in out.items()}
    c = w.contacts()
    res["contact_ab"] = np.stack([c["a"], c["b"]], axis=1)
    res["contact_lambda"] = c["lambda"]
    res["contact_penalty"] = c["penalty"]
    res["pairs"] = w.pairs()
    return res


def write(name: str, scene, steps: int, every: int) -> None:
    data = {"steps": steps, "every": every}
    for prec in (32, 64):
        t = trajectory(scene, steps, every, prec)
        for k, v in t.items():
            data[f"fp{prec}_{k}"] = v.astype(np.float32) if prec == 32 and v.dtype == np.float64 else v
    np.savez_compressed(HERE / name, **data)
    print("wrote", HERE / name)


def write_dtraj(name: str, scene, steps: int, every: int) -> None:
    """The FP32 oracle's trajectory as a scene and trajectory file
    (diomedes_amd.trajectory): frames every `every` steps with velocities and
    the step's broad-phase pairs."""
    from diomedes_amd import trajectory as T
    from diomedes_amd import records as R

    w = pyoracle.OracleWorld(scene, 32)
    with T.Writer(HERE / name, scene, T.FLAG_VELOCITIES | T.FLAG_PAIRS) as out:
        st = w.state()
        out.frame(0, st["positions"], st["orientations"], st["velocities"], st["angular_velocities"],
                  np.zeros((0, 2), np.uint32))
        for k in range(steps // every):
            w.step(nsteps=every)
            st = w.state()
            pr = w.pairs()
            out.frame((k + 1) * every, st["positions"], st["orientations"], st["velocities"],
                      st["angular_velocities"], pr)
    print("wrote", HERE / name)


if __name__ == "__main__":
    pyoracle.threads(1)
    write("box_stack_s0.npz", scenes.box_stack(0), 300, 10)
    write("cube_field_8x2x8_s7.npz", scenes.cube_field(7, nx=8, ny=2, nz=8), 120, 20)
    write_dtraj("box_stack_s0.dtraj", scenes.box_stack(0), 300, 10)

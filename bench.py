"""Benchmark: AVBD body-iterations/s and ms/step on the 8192-cube field.

Workload: 32 x 32 x 8 = 8192 cubes (h ~ U(0.25, 0.75), density 1, spacing 2.0,
+-0.1 jitter, random orientations, seed 0) dropped onto the floor, dt = 1/60,
4 AVBD iterations, friction 0.6, FP32, the reference's step protocol
(contact lag 1).  All data is synthetic and resident in HBM before timing.

One "step" = broad phase + narrow phase + contact build + CSR + AVBD
(init, 4 x (primal, dual), recover), fully on the device.

value = N_dyn * I * K / t_AVBD, where t_AVBD is the summed GPU time of the
AVBD segment (HIP events around it in every step).  ms_per_step is the wall
time of the whole K-step timed region / K.  The roofline object describes the
primal kernel (the dominant one), timed per launch with HIP events in a
profiled pass over the same workload.  cpu_baseline is the FP32 oracle (the
C++ restatement of the reference algorithms, OpenMP over the host cores)
stepping the same scene for the same steps on the host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def avbd_bytes(n_dyn: int, n: int, iters: int, m: float, k: float) -> float:
    """Algorithmic (compulsory) HBM bytes of one AVBD solve: init 152 N_dyn;
    I primal launches of 136 N_dyn + 8 K + 100 M, plus 24 M for the fused
    dual update's lambda / penalty store in iterations 1 .. I-1; the final
    dual + recover 124 M + 56 N + 92 N_dyn."""
    return (152 * n_dyn + iters * (136 * n_dyn + 8 * k + 100 * m) + (iters - 1) * 24 * m
            + 124 * m + 56 * n + 92 * n_dyn)


def primal_bytes(n_dyn: int, m: float, k: float, iters: int) -> float:
    """Primal kernel, average over the I launches of a solve: 136 B per
    dynamic body (state, snapshot, initial, inertial, CSR offsets, new pose),
    8 B per CSR entry (contact index, peer code), 100 B per contact (its
    record's used lanes) and, in I - 1 of the I launches, the 24 B lambda /
    penalty store of the fused dual update.  Gathers of neighbour poses are
    counted once per body (they hit cache)."""
    return 136 * n_dyn + 8 * k + 100 * m + 24 * m * (iters - 1) / iters


PROFILE_MD = ROOT / "profiles" / "cube_field_8192_kernels.md"


def measured_traffic(kernel: str) -> dict | None:
    """HBM bytes per primal launch from the committed rocprofv3 counter runs
    (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy kernel), when they were
    taken on the kernel this run uses."""
    try:
        for line in PROFILE_MD.read_text().splitlines():
            if line.startswith("primal_traffic_json:"):
                t = json.loads(line.split(":", 1)[1])
                return t if t.get("kernel") == kernel else None
    except (OSError, ValueError):
        return None
    return None


def batched_leg(args, tuning: dict) -> dict:
    """The batched-scenes workload: every scene of the table stepped by the
    same launches (one world with a scene table, the step captured in HIP
    graphs), against a sample of the same scenes stepped one plain context
    each, one after the other.  Both report AVBD body-iters/s (HIP events
    around the AVBD segment) and full-step body-iters/s (wall clock)."""
    from diomedes_amd import PhysicsContext, scenes

    batch = scenes.batched_scenes(args.seed, args.batch_scenes)
    opts = batch.options
    iters = opts.iterations
    k = args.steps
    n_dyn = batch.n_dynamic
    with PhysicsContext(batch.n, max_pairs=8 * batch.n, max_contacts=16 * batch.n,
                        max_hull_vertices=len(batch.hull)) as ctx:
        ctx.set_tuning(**tuning)
        ctx.batch_upload(batch)
        if args.warmup:
            ctx.step(opts, args.warmup)
        ctx.step(opts, 0)
        t0 = time.perf_counter()
        ctx.step(opts, k)
        wall = time.perf_counter() - t0
        st = ctx.stats()
    _progress(f"batched leg: {len(batch.scenes)} scenes, {k} steps in {wall:.2f} s")
    out = {
        "workload": f"{len(batch.scenes)} scenes, sizes round(exp(U(ln 8, ln 512))), unit cubes and r 0.5 "
                    f"icosphere hulls dropped on per-scene floors, I={iters}, dt=1/60, friction 0.5, seed {args.seed}",
        "bodies": batch.n,
        "value": n_dyn * iters * k / (st.ms_avbd / 1000.0),
        "unit": "body-iters/s",
        "ms_per_step": 1000.0 * wall / k,
        "ms_per_step_avbd": st.ms_avbd / k,
        "body_iters_per_s_full_step": n_dyn * iters * k / wall,
        "last_step": {"pairs": st.pairs, "contacts": st.contacts, "csr_nnz": st.csr_nnz},
    }
    if args.plain_sample <= 0:
        return out
    # plain: a sample of the same scenes, each in its own context, stepped in turn
    step_ = max(1, len(batch.scenes) // max(1, args.plain_sample))
    sample = list(range(0, len(batch.scenes), step_))[: args.plain_sample]
    bi = 0.0
    t_avbd = 0.0
    t_wall = 0.0
    for s_i in sample:
        sc = batch.scenes[s_i]
        with PhysicsContext(sc.n, max_hull_vertices=len(sc.hull)) as one:
            one.set_tuning(**tuning)
            one.upload(sc)
            if args.warmup:
                one.step(opts, args.warmup)
            one.step(opts, 0)
            t0 = time.perf_counter()
            one.step(opts, k)
            t_wall += time.perf_counter() - t0
            t_avbd += one.stats().ms_avbd / 1000.0
        bi += sc.n_dynamic * iters * k
    _progress(f"plain sample: {len(sample)} scenes in {t_wall:.2f} s")
    plain_avbd = bi / t_avbd
    plain_full = bi / t_wall
    out["plain"] = {
        "sample": f"every {step_}th scene of the table ({len(sample)} scenes, "
                  f"{sum(batch.scenes[i].n for i in sample)} bodies), one context each, {k} steps each",
        "value": plain_avbd,
        "body_iters_per_s_full_step": plain_full,
    }
    out["batched_over_plain"] = out["value"] / plain_avbd
    out["batched_over_plain_full_step"] = out["body_iters_per_s_full_step"] / plain_full
    return out


def coloured_leg(args, tuning: dict) -> dict:
    """The same cube field stepped in the coloured Gauss-Seidel mode (dio_step
    mode 1): the step's body graph is coloured on the device and each
    iteration runs the primal colour by colour.  t_AVBD here includes the
    colouring and its one host read-back per step."""
    from diomedes_amd import PhysicsContext, scenes

    scene = scenes.cube_field(args.seed)
    opts = scene.options
    k = args.steps
    with PhysicsContext(scene.n) as ctx:
        ctx.set_tuning(**tuning)
        ctx.upload(scene)
        if args.warmup:
            ctx.step(opts, args.warmup, mode=1)
        ctx.step(opts, 0)
        t0 = time.perf_counter()
        ctx.step(opts, k, mode=1)
        wall = time.perf_counter() - t0
        st = ctx.stats()
    _progress(f"coloured leg: {k} steps in {wall:.2f} s")
    return {
        "workload": "the same cube field, coloured Gauss-Seidel primal (mode 1)",
        "value": scene.n_dynamic * opts.iterations * k / (st.ms_avbd / 1000.0),
        "unit": "body-iters/s",
        "ms_per_step": 1000.0 * wall / k,
        "ms_per_step_avbd": st.ms_avbd / k,
        "last_step": {"colours": st.colours, "colour_rounds": st.colour_rounds, "contacts": st.contacts},
    }


def overlap_leg(args, tuning: dict) -> dict:
    """The same cube field with the next step's broad and narrow phases on a
    second stream beside the contact build and CSR (tuning overlap_broad):
    a shorter step, while the second hardware queue slows the dispatch of
    the AVBD launches that follow."""
    from diomedes_amd import PhysicsContext, scenes

    scene = scenes.cube_field(args.seed)
    opts = scene.options
    k = args.steps
    with PhysicsContext(scene.n) as ctx:
        ctx.set_tuning(**tuning)
        ctx.set_tuning(overlap_broad=1)
        ctx.upload(scene)
        if args.warmup:
            ctx.step(opts, args.warmup)
        ctx.step(opts, 0)
        t0 = time.perf_counter()
        ctx.step(opts, k)
        wall = time.perf_counter() - t0
        st = ctx.stats()
    _progress(f"overlap leg: {k} steps in {wall:.2f} s")
    return {
        "workload": "the same cube field, broad + narrow overlapped with the contact build (overlap_broad = 1)",
        "value": scene.n_dynamic * opts.iterations * k / (st.ms_avbd / 1000.0),
        "unit": "body-iters/s",
        "ms_per_step": 1000.0 * wall / k,
        "ms_per_step_avbd": st.ms_avbd / k,
    }


def mixed_leg(args, tuning: dict) -> dict:
    """The mixed workload: 20000 hulls (42-vertex icospheres), tetrahedra and
    boxes with N(0, 2) velocities in a 60 m container, 128 ropes of 64 links and
    512 springs; GJK + EPA for the convex pairs, dt = 1/120, 8 iterations."""
    from diomedes_amd import PhysicsContext, scenes

    scene = scenes.mixed_field(args.seed)
    opts = scene.options
    k = args.mixed_steps
    with PhysicsContext(scene.n, max_hull_vertices=len(scene.hull), max_constraints=len(scene.constraints)) as ctx:
        ctx.set_tuning(**tuning)
        ctx.upload(scene)
        if args.warmup:
            ctx.step(opts, args.warmup)
        ctx.step(opts, 0)
        t0 = time.perf_counter()
        ctx.step(opts, k)
        wall = time.perf_counter() - t0
        st = ctx.stats()
    _progress(f"mixed leg: {k} steps in {wall:.2f} s")
    return {
        "workload": f"mixed field: 20000 convex bodies + 8192 rope links, {len(scene.constraints)} joints, "
                    f"I={opts.iterations}, dt=1/120, {k} steps after {args.warmup}",
        "bodies": scene.n,
        "value": scene.n_dynamic * opts.iterations * k / (st.ms_avbd / 1000.0),
        "unit": "body-iters/s",
        "ms_per_step": 1000.0 * wall / k,
        "ms_per_step_avbd": st.ms_avbd / k,
        "last_step": {"pairs": st.pairs, "contacts": st.contacts, "narrow_contacts": st.narrow_contacts},
    }


def _joint_error(scene, pos, rot) -> float:
    """Worst violation over the scene's positional joints: |p_a - p_b| for a
    socket, ||delta| - L| for a rod, max(|delta| - L, 0) for a rope (metres;
    bending joints, whose anchor field holds a rest rotation, are left out)."""
    import numpy as np

    c = scene.constraints
    c = c[(c["kind"] != 7) & (c["kind"] != 8)]

    def world(idx, anchor):
        q = rot[idx].astype(np.float64)
        u, w = q[:, :3], q[:, 3:4]
        v = anchor.astype(np.float64)
        return pos[idx].astype(np.float64) + v + 2.0 * np.cross(u, np.cross(u, v) + w * v)

    delta = world(c["a"], c["anchor_a"][:, :3]) - world(c["b"], c["anchor_b"][:, :3])
    d = np.linalg.norm(delta, axis=1)
    length = c["anchor_a"][:, 3].astype(np.float64)
    err = np.where(c["kind"] == 5, d, np.where(c["kind"] == 6, np.maximum(d - length, 0.0), np.abs(d - length)))
    return float(err.max()) if len(err) else 0.0


def joints_leg(args, tuning: dict) -> dict:
    """Hard joints: 256 chains of 32 cubes (8,448 bodies, 8,192 hard sockets,
    rods and ropes) released horizontally from static anchors, swinging into
    each other and piling on the floor; coloured mode (mode 1, which hard
    joints between dynamic bodies need), I = 8, dt = 1/60."""
    from diomedes_amd import PhysicsContext, scenes

    scene = scenes.hard_chains(args.seed)
    opts = scene.options
    k = args.joint_steps
    with PhysicsContext(scene.n, max_constraints=len(scene.constraints)) as ctx:
        ctx.set_tuning(**tuning)
        ctx.upload(scene)
        if args.warmup:
            ctx.step(opts, args.warmup, mode=1)
        ctx.step(opts, 0, mode=1)
        t0 = time.perf_counter()
        ctx.step(opts, k, mode=1)
        wall = time.perf_counter() - t0
        st = ctx.stats()
        d = ctx.download()
    err = _joint_error(scene, d["positions"], d["orientations"])
    _progress(f"joints leg: {k} steps in {wall:.2f} s, worst joint error {err:.3e} m")
    return {
        "workload": f"hard chains: 256 x 32 cubes on {len(scene.constraints)} hard joints, coloured mode, "
                    f"I={opts.iterations}, dt=1/60, {k} steps after {args.warmup}",
        "bodies": scene.n,
        "value": scene.n_dynamic * opts.iterations * k / (st.ms_avbd / 1000.0),
        "unit": "body-iters/s",
        "ms_per_step": 1000.0 * wall / k,
        "ms_per_step_avbd": st.ms_avbd / k,
        "worst_joint_error_m": err,
        "last_step": {"colours": st.colours, "colour_rounds": st.colour_rounds, "contacts": st.contacts},
    }


def _wall_loss_grad(pos, targets, offsets):
    """Sum over seeds of the mean squared distance of the bricks' final
    centres to their targets, and its adjoint with respect to the final
    positions (float64)."""
    import numpy as np

    loss = 0.0
    d = np.zeros(pos.shape, np.float64)
    for s, tgt in enumerate(targets):
        lo = int(offsets[s])
        nb = len(tgt)
        diff = pos[lo:lo + nb].astype(np.float64) - tgt
        loss += float(np.mean(np.sum(diff * diff, axis=1)))
        d[lo:lo + nb] = 2.0 * diff / nb
    return loss, d


def adjoint_leg(args) -> dict:
    """The differentiable wall throw: args.adjoint_seeds walls of 64 x 32
    bricks, each hit by its own sphere, batched in one world and stepped
    args.adjoint_steps steps with the adjoint recording on (checkpoint every
    8 steps); then one reverse-mode pass gives d loss / d (each projectile's
    v0 and mass, the friction).  forward_ms is the recorded forward (the full
    step: broad, narrow, contacts, AVBD, plus the recording); backward_ms the
    reverse pass (each segment replayed once from its checkpoint with its
    iterations taped, then reversed); replay_ms a solve-only replay of the
    recording.  The reverse-mode gradient of seed 0 over 20 steps is reported
    as a figure; its check is the test suite's (reverse mode against the
    oracle's float-valued dual numbers, the exact derivative of the float32
    replay's piece): float32 central differences at any usable step cross
    pieces of the piecewise-smooth float32 map and measure rounding, not the
    derivative."""
    import numpy as np

    from diomedes_amd import PhysicsContext, scenes

    seeds, steps, every = args.adjoint_seeds, args.adjoint_steps, 8
    throws = [scenes.wall_throw(args.seed + s) for s in range(seeds)]
    batch = scenes.BatchedScenes(throws, np.zeros((0, 3), np.float32), throws[0].options)
    opts = throws[0].options
    targets = [scenes.wall_throw_target(sc, args.seed + s) for s, sc in enumerate(throws)]
    offs = batch.offsets

    def grad5(g, state, s):
        i = int(offs[s]) + throws[s].n - 1
        inv_m = float(state["inv_mass"][i])
        c = state["inv_moment"][i, :3].astype(np.float64) / inv_m
        m = 1.0 / inv_m
        gm = -(float(g["inv_mass"][i]) + float(np.dot(g["inv_moment"][i].astype(np.float64), c))) / (m * m)
        return [*map(float, g["velocities"][i]), gm, float(g["friction"])]

    n = batch.n
    with PhysicsContext(n, max_pairs=8 * n, max_contacts=16 * n) as ctx:
        ctx.batch_upload(batch)
        ctx.step(opts, 3)  # first launches and graph captures outside the timing
        ctx.batch_upload(batch)
        # about 6 contacts per body per step on average over the throw (11 at most)
        ctx.adjoint_begin(steps, every, max_recorded_contacts=8 * n * steps)
        ctx.step(opts, 0)
        t0 = time.perf_counter()
        ctx.step(opts, steps)
        fwd = time.perf_counter() - t0
        _progress(f"adjoint leg: recorded forward of {steps} steps, {n} bodies, {fwd:.2f} s")
        pos = ctx.download()["positions"]
        loss, dpos = _wall_loss_grad(pos, targets, offs)
        ctx.adjoint_backward(opts, d_positions=dpos.astype(np.float32))  # first call sizes the scratch
        t0 = time.perf_counter()
        g = ctx.adjoint_backward(opts, d_positions=dpos.astype(np.float32))
        bwd = time.perf_counter() - t0
        t0 = time.perf_counter()
        rpos, _ = ctx.adjoint_replay(opts)
        rep = time.perf_counter() - t0
        _progress(f"adjoint leg: backward {bwd:.2f} s, replay {rep:.2f} s")
        replay_exact = bool(np.array_equal(rpos.view(np.uint32), pos.view(np.uint32)))
    # the gradient of one seed over 20 steps, where the tangent map is still well conditioned
    short = 20
    sc0 = throws[0]
    with PhysicsContext(sc0.n) as c1:
        c1.upload(sc0)
        c1.adjoint_begin(short, every)
        c1.step(opts, short)
        p1 = c1.download()["positions"]
        _, d1 = _wall_loss_grad(p1, targets[:1], offs[:1])
        g1 = c1.adjoint_backward(opts, d_positions=d1.astype(np.float32))
        g_short = grad5(g1, sc0.state, 0)
    bodies_dyn = batch.n_dynamic
    return {
        "workload": f"{seeds} wall throws (64 x 32 bricks 0.5 x 0.25 x 0.25 + one r 0.5, m 20 sphere each) "
                    f"batched, {steps} steps, I={opts.iterations}, dt=1/60, checkpoint every {every}, seeds "
                    f"{args.seed}..{args.seed + seeds - 1}",
        "bodies": n,
        "forward_ms": 1000.0 * fwd,
        "backward_ms": 1000.0 * bwd,
        "replay_ms": 1000.0 * rep,
        "backward_over_forward": bwd / fwd,
        "backward_over_replay": bwd / rep,
        "body_iters_per_s_backward": bodies_dyn * opts.iterations * steps / bwd,
        "replay_bit_exact": replay_exact,
        "loss": loss,
        "grad_seed0_20_steps": dict(zip(("v0x", "v0y", "v0z", "mass", "friction"), g_short)),
        "note": "the 120-step backward is timed but its gradient is not reported: the penalty ramp makes the "
                "tangent map expansive (FP64: max |dx/dv0| 3e9 by 60 steps) and the float32 tangents of 120 steps "
                "overflow. The 20-step gradient equals the oracle's float-valued dual numbers "
                "(tests/test_gpu_adjoint.py); float32 central differences are no check of it, since at any usable "
                "step they cross pieces of the piecewise-smooth float32 map",
    }


def _progress(msg: str) -> None:
    """Progress lines go to stderr; stdout carries only the JSON result."""
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def cpu_baseline(scene, warmup: int, steps: int, budget_s: float = 25.0, mode: int = 0) -> dict:
    """FP32 oracle on the host cores over the same warm-up and steps (mode 1:
    its coloured Gauss-Seidel solve)."""
    from oracle import pyoracle

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    pyoracle.threads(cores)
    w = pyoracle.OracleWorld(scene, 32)
    w.set_mode(mode)
    w.step(nsteps=warmup)
    w.reset_times()
    done = 0
    t0 = time.perf_counter()
    while done < steps and time.perf_counter() - t0 < budget_s:
        w.step(nsteps=1)
        done += 1
    wall = time.perf_counter() - t0
    t = w.times()
    n_dyn = scene.n_dynamic
    iters = scene.options.iterations
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": n_dyn * iters * done / t["avbd"] if t["avbd"] > 0 else None,
        "unit": "AVBD body-iters/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{scene.name}: FP32 oracle, steps {warmup}..{warmup + done} after {warmup} warm-up steps, "
                  f"{cores} OpenMP threads on {cpu_model or 'the host CPU'}",
        "ms_per_step": 1000.0 * wall / max(done, 1),
        "ms_per_step_avbd": 1000.0 * t["avbd"] / max(done, 1),
        "ms_per_step_broad": 1000.0 * t["broad"] / max(done, 1),
        "steps_timed": done,
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--batch-scenes", type=int, default=4096,
                    help="scenes of the batched leg (0: skip it); steps and warm-up as for the main workload")
    ap.add_argument("--plain-sample", type=int, default=64,
                    help="scenes of the batch stepped one context each for the plain comparison")
    ap.add_argument("--profile-steps", type=int, default=20, help="steps of the per-kernel timing pass")
    ap.add_argument("--tune-steps", type=int, default=0,
                    help="steps of a separate run of the same scene after which dio_autotune picks the launch "
                         "shapes and dual schedule (0: library defaults)")
    ap.add_argument("--no-coloured", action="store_true", help="skip the coloured Gauss-Seidel leg")
    ap.add_argument("--no-overlap", action="store_true", help="skip the overlapped broad-phase leg")
    ap.add_argument("--mixed-steps", type=int, default=300, help="steps of the mixed-field leg (0: skip it)")
    ap.add_argument("--joint-steps", type=int, default=120, help="steps of the hard-joints leg (0: skip it)")
    ap.add_argument("--only-leg", choices=("coloured", "batched", "adjoint", "mixed", "joints", "overlap"),
                    default=None,
                    help="run only this leg and print its object (for profiling it alone)")
    ap.add_argument("--adjoint-seeds", type=int, default=64,
                    help="wall throws of the adjoint leg (0: skip it)")
    ap.add_argument("--adjoint-steps", type=int, default=120, help="steps of the adjoint leg")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (reference for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py: --gpus {args.gpus} is not supported: the step runs on one MI355X. A scene table would "
            "split across GPUs by scene with no exchange; one large scene would split by Morton range with a "
            "pose halo per iteration. Neither is run here.",
            file=sys.stderr,
        )
        return 2

    from diomedes_amd import PhysicsContext, scenes

    if args.only_leg:
        legs = {"coloured": lambda: coloured_leg(args, {}), "batched": lambda: batched_leg(args, {}),
                "adjoint": lambda: adjoint_leg(args), "mixed": lambda: mixed_leg(args, {}),
                "joints": lambda: joints_leg(args, {}), "overlap": lambda: overlap_leg(args, {})}
        print(json.dumps({"leg": args.only_leg, **legs[args.only_leg]()}), flush=True)
        return 0

    if args.calibrate_copy_mib:
        import torch

        n = args.calibrate_copy_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    scene = scenes.cube_field(args.seed)
    opts = scene.options
    n, n_dyn, iters = scene.n, scene.n_dynamic, opts.iterations

    tuning = None
    if args.tune_steps:
        # launch shapes chosen on the settled scene in a separate context
        with PhysicsContext(n) as tctx:
            tctx.upload(scene)
            tctx.step(opts, args.tune_steps)
            tctx.autotune(opts, reps=20)
            tuning = tctx.tuning()

    with PhysicsContext(n) as ctx:
        if tuning:
            ctx.set_tuning(**tuning)
        tuning = ctx.tuning()
        ctx.upload(scene)
        if args.warmup:
            ctx.step(opts, args.warmup)
        _progress(f"warm-up done ({args.warmup} steps)")
        ctx.step(opts, 0)  # synchronise (a zero-step call syncs the stream)
        t0 = time.perf_counter()
        ctx.step(opts, args.steps)  # returns after the stream is synchronised
        _progress(f"timed region done ({args.steps} steps)")
        wall = time.perf_counter() - t0
        st = ctx.stats()
        ms_avbd = st.ms_avbd
        sum_m, sum_k, sum_p = st.sum_contacts, st.sum_csr_nnz, st.sum_pairs
        last = {"pairs": st.pairs, "contacts": st.contacts, "csr_nnz": st.csr_nnz}
        # per-kernel timing pass (eager launches, HIP events around each AVBD kernel)
        ctx.step(opts, args.profile_steps, profile=True)
        _progress(f"profile pass done ({args.profile_steps} steps)")
        pst = ctx.stats()

    k = args.steps
    value = n_dyn * iters * k / (ms_avbd / 1000.0)
    avg_m = sum_m / k
    avg_k = sum_k / k
    # roofline of the primal kernel
    p_launches = max(pst.n_primal, 1)
    primal_ms = pst.ms_primal / p_launches
    pm = pst.sum_contacts / max(pst.steps, 1)
    pk = pst.sum_csr_nnz / max(pst.steps, 1)
    pbytes = primal_bytes(n_dyn, pm, pk, iters)
    achieved = pbytes / (primal_ms / 1000.0) / 1e9
    step_bytes = avbd_bytes(n_dyn, n, iters, avg_m, avg_k)
    lanes = tuning["primal_lanes"]
    # the step's primal launches: with the fused schedule the first of a step is
    # <G, false> and the others <G, true>; the roofline averages over all of them
    kname = (f"k_avbd_primal_grouped<{lanes}, {'*' if tuning['fuse_dual'] else 'false'}>" if lanes > 1
             else f"k_avbd_primal<{'*' if tuning['fuse_dual'] else 'false'}>")
    traffic = measured_traffic(kname)
    line = {
        "metric": "AVBD body-iters/s & ms/step, 8k boxes",
        "value": value,
        "unit": "body-iters/s",
        "n_gpus": 1,
        "steps": k,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall / k,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded cube field)",
        "config": {
            "workload": "plain cube field 32x32x8 = 8192 boxes, I=4, dt=1/60, friction 0.6, contact lag 1",
            "bodies": n,
            "iterations": iters,
            "seed": args.seed,
            "mode": "jacobi (reference parity)",
        },
        "tuning": tuning,
        "ms_per_step_avbd": ms_avbd / k,
        "body_iters_per_s_full_step": n_dyn * iters * k / wall,
        "avg_contacts": avg_m,
        "avg_csr_nnz": avg_k,
        "avg_pairs": sum_p / k,
        "last_step": last,
        "avbd_segment_gbs": step_bytes * k / (ms_avbd / 1000.0) / 1e9,
        "roofline": {
            "kernel": kname,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic["traffic_bytes"] if traffic else None,
            "traffic_source": (f"{PROFILE_MD.relative_to(ROOT)}: rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, "
                               f"{traffic['window']}") if traffic else None,
            "avg_launch_us": primal_ms * 1000.0,
            "bytes_per_launch": pbytes,
            "launches_timed": pst.n_primal,
        },
        "kernel_avg_us": {
            "init": 1000.0 * pst.ms_init / max(pst.n_init, 1),
            "primal": 1000.0 * primal_ms,
            "dual_recover": 1000.0 * pst.ms_recover / max(pst.n_recover, 1),
        },
    }
    if not args.no_overlap:
        line["overlap"] = overlap_leg(args, tuning)
    if not args.no_coloured:
        line["coloured"] = coloured_leg(args, tuning)
    if args.mixed_steps:
        line["mixed"] = mixed_leg(args, tuning)
    if args.joint_steps:
        line["joints"] = joints_leg(args, tuning)
    if args.batch_scenes:
        line["batched"] = batched_leg(args, tuning)
    if args.adjoint_seeds:
        line["adjoint"] = adjoint_leg(args)
    if not args.no_cpu:
        cb = cpu_baseline(scenes.cube_field(args.seed), args.warmup, k)
        line["cpu_baseline"] = cb
        if cb.get("value"):
            line["gpu_over_cpu"] = value / cb["value"]
        if args.joint_steps:
            jc = cpu_baseline(scenes.hard_chains(args.seed), args.warmup, args.joint_steps, budget_s=10.0, mode=1)
            line["joints"]["cpu_baseline"] = jc
            if jc.get("value"):
                line["joints"]["gpu_over_cpu"] = line["joints"]["value"] / jc["value"]
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

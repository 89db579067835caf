"""The oracle's pins, checked against planted mistakes.

Each entry below plants one plausible transcription mistake in a copy of the
oracle's sources (a sign, a dropped term or vertex, a swapped weight or
index, a reversed tie rule), builds that copy, and runs the designed pin
tests named for it against the mutant (ORACLE_LIB points the ctypes wrapper
at it).  The entry passes when at least one of those tests fails.  The
unmutated sources, built the same way, must pass every named test.

The named tests are designed checks only (the reference's own unit tests,
closed forms, library routines, brute force, invariants); the golden
fixtures and the adjoint tests are left out, since they would catch any
change without saying what was wrong.  Each snippet must occur exactly once
in its file, so an edit of the oracle that moves it fails here loudly.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
ORACLE = ROOT / "oracle"
HPP, CAPI = "avbd_oracle.hpp", "oracle_capi.cpp"
CXXFLAGS = ["-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-march=x86-64-v2",
            "-shared"]

P, K, G, J, C, I = ("tests/test_oracle_pins.py", "tests/test_oracle_reference_kats.py",
                    "tests/test_oracle_geometry_pins.py", "tests/test_oracle_joints.py",
                    "tests/test_oracle_colouring.py", "tests/test_oracle_instances.py")

# (name, file, snippet, replacement, tests that must catch it)
MUTATIONS = [
    # ---- narrow phase: support, SAT manifold, EPA
    ("hull support skips the last vertex", HPP,
     "for (uint32_t i = 1; i < count; ++i) {\n                T dd = dot(hull[first + i], d);",
     "for (uint32_t i = 1; i + 1 < count; ++i) {\n                T dd = dot(hull[first + i], d);",
     [G + "::test_hull_support_is_the_first_maximal_vertex", G + "::test_hull_aabb_is_the_extent_of_the_rotated_vertices"]),
    ("kernel tetrahedron support keeps the last maximal corner", HPP,
     "if (dd > best_dot) { best_dot = dd; best = c; }", "if (dd >= best_dot) { best_dot = dd; best = c; }",
     [K + "::test_tetrahedron_support_is_the_farthest_corner"]),
    ("host tetrahedron support keeps the first maximal corner", HPP,
     "if (k >= best_key) { best_key = k; best = i; }", "if (k > best_key) { best_key = k; best = i; }",
     [K + "::test_tetrahedron_support_is_the_farthest_corner"]),
    ("SAT reference point not projected onto the reference face", HPP,
     "V3<T> rp = ip + ref_n * rmax(pd, T(0));", "V3<T> rp = ip;",
     [G + "::test_sat_manifold_points", G + "::test_sat_resting_cube_closed_form"]),
    ("clipping interpolates with 1 - t", HPP,
     "out[count++] = prev + (cur - prev) * t;", "out[count++] = prev + (cur - prev) * (T(1) - t);",
     [G + "::test_sat_manifold_points", G + "::test_sat_resting_cube_closed_form"]),
    ("SAT edge axes skipped", HPP,
     "if (!test_axis(cross(aa[i], ba[j]), 2, i * 3 + j, a, aa, b, ba, delta, best)) return false;",
     "if (false && !test_axis(cross(aa[i], ba[j]), 2, i * 3 + j, a, aa, b, ba, delta, best)) return false;",
     [P + "::test_box_sat_depth_and_normal_match_brute_force", G + "::test_sat_manifold_points"]),
    ("box radius repeats the second axis", HPP,
     "b.corners[0][2] * ofabs(dot(axis, ax[2]));", "b.corners[0][2] * ofabs(dot(axis, ax[1]));",
     [P + "::test_box_sat_depth_and_normal_match_brute_force", G + "::test_sat_manifold_points"]),
    ("incident axis chosen without the absolute value", HPP,
     "T d = ofabs(dot(ax[i], n));", "T d = dot(ax[i], n);",
     [G + "::test_sat_manifold_points"]),
    ("face frame corner 1 with the wrong signs", HPP,
     "verts[1] = center - u * eu + v * ev;", "verts[1] = center + u * eu - v * ev;",
     [G + "::test_sat_manifold_points", G + "::test_sat_resting_cube_closed_form"]),
    ("centred contact point_b on the wrong side", HPP,
     "m.point_b[0] = b.position - n * rb;", "m.point_b[0] = b.position + n * rb;",
     [G + "::test_sat_manifold_points"]),
    ("sphere-sphere normal reversed", HPP,
     "V3<T> n = delta / distance;", "V3<T> n = -delta / distance;",
     [P + "::test_sphere_sphere_closed_form"]),
    ("EPA barycentric weights of the witnesses swapped", HPP,
     "V3<T> pa = wa * w1 + wb * w2 + wc * w3;", "V3<T> pa = wa * w1 + wb * w3 + wc * w2;",
     [G + "::test_epa_witness_points"]),
    ("EPA point_b adds the Minkowski point", HPP,
     "return Penetration<T>{f.normal, f.distance, pa, pa - mk};",
     "return Penetration<T>{f.normal, f.distance, pa, pa + mk};",
     [G + "::test_epa_witness_points", G + "::test_epa_sphere_resting_on_a_box"]),
    ("EPA stops at a loose tolerance", HPP,
     "if (sd - mf.distance < T(1e-4))", "if (sd - mf.distance < T(1e-1))",
     [P + "::test_epa_depth_matches_minkowski_hull", G + "::test_epa_depth_of_hulls_matches_minkowski_hull",
      G + "::test_epa_sphere_resting_on_a_box"]),
    # ---- contact construction
    ("tangents from the negated normal", HPP,
     "    tangents(normal, t1, t2);\n    c.a = a;", "    tangents(-normal, t1, t2);\n    c.a = a;",
     [P + "::test_raw_contact_frame_and_lever_arms"]),
    ("tangent branch tests the wrong component", HPP,
     "ofabs(n.x) > T(0.9) ? cross(n, V3<T>{T(0), T(1), T(0)})",
     "ofabs(n.y) > T(0.9) ? cross(n, V3<T>{T(0), T(1), T(0)})",
     [G + "::test_tangent_frames_on_the_axes"]),
    ("contact c0 sign", HPP,
     "    c.c0 = V3<T>{o.margin - depth, T(0), T(0)};\n    c.penalty",
     "    c.c0 = V3<T>{depth - o.margin, T(0), T(0)};\n    c.penalty",
     [P + "::test_raw_contact_frame_and_lever_arms"]),
    ("container contacts ignore the predicted depth", HPP,
     "T depth = rmax(cur, pd);", "T depth = cur;",
     [P + "::test_container_walls_closed_form"]),
    ("warm start takes the first old contact of a key", HPP,
     "c.lambda = old_contacts[last->second].lambda;\n                c.penalty = old_contacts[last->second].penalty;",
     "c.lambda = old_contacts[keyed[q].second].lambda;\n                c.penalty = old_contacts[keyed[q].second].penalty;",
     [P + "::test_warm_start_and_csr_match_hashmap_semantics"]),
    ("warm start scales lambda by alpha only", HPP,
     "T ag = o.alpha * o.gamma;\n        c.lambda = c.lambda * ag;",
     "T ag = o.alpha;\n        c.lambda = c.lambda * ag;",
     [P + "::test_warm_start_and_csr_match_hashmap_semantics"]),
    ("CSR leaves out body b's entries", HPP,
     "if (contacts[i].b != CONTAINER) indices[cursor[contacts[i].b]++] = i;",
     "if (false) indices[cursor[contacts[i].b]++] = i;",
     [P + "::test_warm_start_and_csr_match_hashmap_semantics"]),
    # ---- the AVBD solve
    ("solve6 l42 reads l31", HPP,
     "T l42 = (a42 - l21 * l41 * d1) / d2;", "T l42 = (a42 - l21 * l31 * d1) / d2;",
     [P + "::test_solve6_matches_numpy"]),
    ("solve6 back substitution drops a term", HPP,
     "T x2 = z2 - l32 * x3 - l42 * x4 - l52 * x5 - l62 * x6;", "T x2 = z2 - l32 * x3 - l42 * x4 - l52 * x5;",
     [P + "::test_solve6_matches_numpy"]),
    ("inertial target without gravity", HPP,
     "if (S.inv_mass[i] > T(0)) ip = ip + pc.gravity * (dt * dt);", "if (false) ip = ip + pc.gravity * (dt * dt);",
     [P + "::test_free_fall_is_semi_implicit_euler", P + "::test_adaptive_gravity_warm_start"]),
    ("adaptive gravity weight not clamped", HPP,
     "T weight = gm > T(0) ? rclamp(dot(accel, g / gm) / gm, T(0), T(1)) : T(0);",
     "T weight = gm > T(0) ? dot(accel, g / gm) / gm : T(0);",
     [P + "::test_adaptive_gravity_warm_start"]),
    ("quaternion step without the half", HPP,
     "return qnormalize(qadd(q, qscale(qmul(p, q), T(0.5))));", "return qnormalize(qadd(q, qscale(qmul(p, q), T(1))));",
     [P + "::test_torque_free_spin_decay"]),
    ("quaternion step multiplies on the other side", HPP,
     "return qnormalize(qadd(q, qscale(qmul(p, q), T(0.5))));", "return qnormalize(qadd(q, qscale(qmul(q, p), T(0.5))));",
     [P + "::test_torque_free_spin_decay"]),
    ("recovered angular velocity without the factor 2", HPP,
     "S.ang[i] = V3<T>{T(2) * rel.x / dt, T(2) * rel.y / dt, T(2) * rel.z / dt};",
     "S.ang[i] = V3<T>{rel.x / dt, rel.y / dt, rel.z / dt};",
     [P + "::test_torque_free_spin_decay"]),
    ("glam rotation with the cross product reversed", HPP,
     "return v * (w * w - b2) + b * (dot(v, b) * T(2)) + cross(b, v) * (w * T(2));",
     "return v * (w * w - b2) + b * (dot(v, b) * T(2)) + cross(v, b) * (w * T(2));",
     [K + "::test_endpoints_apply_body_rotation_and_local_anchors"]),
    ("primal friction bound doubled", HPP,
     "T bounds = ofabs(f0) * ct.friction;\n                T tm", "T bounds = ofabs(f0) * ct.friction * T(2);\n                T tm",
     [P + "::test_sliding_friction_decelerates_at_mu_g"]),
    ("primal normal force clamped to push the wrong way", HPP,
     "T f0 = rmin(ct.penalty.x * c0 + ct.lambda.x, T(0));\n                T f1r",
     "T f0 = rmax(ct.penalty.x * c0 + ct.lambda.x, T(0));\n                T f1r",
     [P + "::test_resting_cube_normal_force_balances_weight"]),
    ("primal drops the other body's linear motion", HPP,
     "c0 = c0 + (dot(r0, dsl) + dot(j0, dsa) - dot(r0, dol) - dot(cross(r_other, r0), doa));",
     "c0 = c0 + (dot(r0, dsl) + dot(j0, dsa) - dot(cross(r_other, r0), doa));",
     [P + "::test_stacked_dynamic_cubes_impulse_balance"]),
    ("primal takes the other body's rotation with the wrong sign", HPP,
     "c0 = c0 + (dot(r0, dsl) + dot(j0, dsa) - dot(r0, dol) - dot(cross(r_other, r0), doa));",
     "c0 = c0 + (dot(r0, dsl) + dot(j0, dsa) - dot(r0, dol) + dot(cross(r_other, r0), doa));",
     [P + "::test_spinning_neighbour_pushes_only_when_its_contact_point_approaches"]),
    ("dual friction bound doubled", HPP,
     "            T bounds = ofabs(f0) * ct.friction;\n            T tm",
     "            T bounds = ofabs(f0) * ct.friction * T(2);\n            T tm",
     [P + "::test_sliding_friction_decelerates_at_mu_g"]),
    ("recovered velocity multiplies by dt", HPP,
     "S.vel[i] = (pos - S.initial_lin[i]) / dt;", "S.vel[i] = (pos - S.initial_lin[i]) * dt;",
     [P + "::test_free_fall_is_semi_implicit_euler"]),
    ("dual penalty ramp by the signed constraint", HPP,
     "if (f0 < T(0)) ct.penalty.x = rmin(ct.penalty.x + pc.beta_lin * ofabs(c0), pc.penalty_max);",
     "if (f0 < T(0)) ct.penalty.x = rmin(ct.penalty.x + pc.beta_lin * c0, pc.penalty_max);",
     [P + "::test_resting_cube_normal_force_balances_weight"]),
    ("spring force sign", HPP,
     "T force = stiff * (len - c.length);", "T force = stiff * (c.length - len);",
     [P + "::test_hanging_spring_equilibrium"]),
    ("socket force sign", HPP,
     "const T fs[3] = {delta.x * stiff, delta.y * stiff, delta.z * stiff};",
     "const T fs[3] = {-delta.x * stiff, -delta.y * stiff, -delta.z * stiff};",
     [P + "::test_ball_socket_holds_the_anchor"]),
    ("soft rope pulls inside its length", HPP,
     "bool active = len > T(1.0e-6) && (c.kind != ROPE || len > c.length);", "bool active = len > T(1.0e-6);",
     [P + "::test_slack_soft_rope_exerts_nothing"]),
    # ---- hard joints and colouring (no reference counterpart)
    ("hard-joint penalty ramp by the signed value", HPP,
     "if (kind != HARD_ROPE || f > T(0)) pen[k] = rmin(pen[k] + beta * ofabs(ch[k]), pmax);",
     "if (kind != HARD_ROPE || f > T(0)) pen[k] = rmin(pen[k] + beta * ch[k], pmax);",
     [J]),
    ("colouring skips the smallest free colour", HPP,
     "while (mask[w] & (1u << b)) ++b;", "while (mask[w] & (1u << b)) ++b;\n                    b += b < 31;",
     [C]),
    # ---- instance transforms (the renderer hand-off)
    ("model matrix x axis with the sign of w z flipped", HPP,
     "const T ax[4] = {T(1) - (yy + zz), xy + wz, xz - wy, T(0)};",
     "const T ax[4] = {T(1) - (yy + zz), xy - wz, xz - wy, T(0)};",
     [I + "::test_rotation_columns_are_the_rotated_basis", I + "::test_rotation_part_is_orthonormal"]),
    ("model matrix without the translation", HPP, "m[12] = pos[0];", "m[12] = T(0);",
     [I + "::test_model_maps_local_vertices_to_world", I + "::test_identity_and_z_rotation_closed_forms"]),
    ("view_proj * model taken as model * view_proj", HPP, "mat4_mul(view_proj, m, out + 32 * (size_t)k);",
     "mat4_mul(m, view_proj, out + 32 * (size_t)k);", [I + "::test_mvp_is_the_matrix_product"]),
    ("instance slots ignore the order", HPP, "const uint32_t b = order ? order[k] : k;", "const uint32_t b = k;",
     [I + "::test_slots_follow_the_order"]),
]


def _plant(dst: Path, mutation) -> None:
    name, fname, old, new, _ = mutation
    for f in (HPP, CAPI):
        text = (ORACLE / f).read_text()
        if f == fname:
            count = text.count(old)
            assert count == 1, f"{name}: the snippet occurs {count} times in oracle/{f}"
            text = text.replace(old, new)
        (dst / f).write_text(text)


def _build(workdir: Path, tag: str, mutation=None) -> Path:
    d = workdir / tag / "oracle"
    d.mkdir(parents=True)
    shutil.copytree(ROOT / "include", workdir / tag / "include")
    if mutation is None:
        for f in (HPP, CAPI):
            shutil.copy(ORACLE / f, d / f)
    else:
        _plant(d, mutation)
    lib = d / "liboracle.so"
    proc = subprocess.run(["g++", *CXXFLAGS, "-o", str(lib), str(d / CAPI)], capture_output=True, text=True)
    assert proc.returncode == 0, f"{tag}: build failed\n{proc.stderr[-2000:]}"
    return lib


def _run(lib: Path, tests) -> subprocess.CompletedProcess:
    env = dict(os.environ, ORACLE_LIB=str(lib), ORACLE_THREADS="2")
    return subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                           *tests], cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)


@pytest.fixture(scope="module")
def workdir(tmp_path_factory):
    return tmp_path_factory.mktemp("oracle_mutants")


def test_mutation_snippets_are_unique():
    for m in MUTATIONS:
        text = (ORACLE / m[1]).read_text()
        assert text.count(m[2]) == 1, f"{m[0]}: snippet occurs {text.count(m[2])} times"


def test_planted_mistakes_fail_the_pins(workdir):
    jobs = max(1, min(8, (len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 4)))
    every_test = sorted({t for m in MUTATIONS for t in m[4]})

    def control():
        lib = _build(workdir, "control")
        return _run(lib, every_test)

    def one(k):
        m = MUTATIONS[k]
        lib = _build(workdir, f"m{k:02d}", m)
        return m[0], _run(lib, m[4])

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        ctl = ex.submit(control)
        results = list(ex.map(one, range(len(MUTATIONS))))
        ctl = ctl.result()
    assert ctl.returncode == 0, f"the unmutated oracle fails its pins:\n{ctl.stdout[-3000:]}"
    survivors = [name for name, r in results if r.returncode != 1]
    # returncode 1: a test failed; 0 would mean every named test passed
    # (the mistake survives); anything else is an error of the run itself
    errors = [(name, r.returncode, r.stdout[-800:]) for name, r in results if r.returncode not in (0, 1)]
    assert not errors, errors
    assert not survivors, f"planted mistakes no designed pin catches: {survivors}"

"""The reference's own unit tests, ported as known-answer tests of the oracle.

The reference holds no golden vectors; these ten #[test]s are its only pins
(engine/src/physics/shape.rs:53-94, constraints.rs:158-186, gpu.rs:421-433,
engine/src/render/gpu_physics.rs:702-721, narrowphase/src/cuboid.rs:404-422).
Each assertion is the reference's, applied to the FP32 oracle (and, where it
runs through the FP64 oracle too, to both).
"""

from __future__ import annotations

import numpy as np
import pytest

from diomedes_amd import records as R
from diomedes_amd.physics import (
    AvbdBody,
    AvbdContainer,
    AvbdOptions,
    BallSocket,
    Rod,
    Rope,
    Shape,
    Spring,
    constraints_from_avbd,
    gpu_bodies,
    state_from_bodies,
)


def _body(shape: Shape, pos=(0.0, 0.0, 0.0), rot=(0.0, 0.0, 0.0, 1.0)):
    b = AvbdBody(shape, tuple(pos), tuple(rot))
    return gpu_bodies([b])[0]


# shape.rs:53-64
def test_sphere_support_tracks_direction(oracle):
    s = _body(Shape.sphere(2.0))
    assert np.array_equal(oracle.support(s, (1.0, 0.0, 0.0), host=True), [2.0, 0.0, 0.0])
    assert np.array_equal(oracle.support(s, (0.0, -3.0, 0.0), host=True), [0.0, -2.0, 0.0])


# shape.rs:66-74
def test_cube_support_is_a_corner(oracle):
    s = _body(Shape.cube(0.5))
    assert np.array_equal(oracle.support(s, (1.0, 1.0, 1.0), host=True), [0.5, 0.5, 0.5])
    assert np.array_equal(oracle.support(s, (-1.0, 1.0, 0.0), host=True), [-0.5, 0.5, 0.5])


# shape.rs:76-94
def test_tetrahedron_support_is_the_farthest_corner(oracle):
    corners = [(1.0, 1.0, 1.0), (1.0, -1.0, -1.0), (-1.0, 1.0, -1.0), (-1.0, -1.0, 1.0)]
    s = _body(Shape.tetrahedron(corners))
    p = oracle.support(s, (1.0, 0.0, 0.0), host=True)
    assert p[0] == 1.0
    assert p @ np.array([1.0, 0.0, 0.0]) == 1.0
    assert np.array_equal(oracle.support(s, (-1.0, -1.0, 0.0), host=True), [-1.0, -1.0, 1.0])
    # the tie along +x: the host Shape::support keeps the last maximal corner
    # (Iterator::max_by), the kernels' support the first (support.rs:44-55)
    assert np.array_equal(p, [1.0, -1.0, -1.0])
    assert np.array_equal(oracle.support(s, (1.0, 0.0, 0.0), host=False), [1.0, 1.0, 1.0])


# constraints.rs:158-174
def test_endpoints_apply_body_rotation_and_local_anchors(oracle):
    half = np.float32(np.float32(90.0) * np.float32(np.pi / 180.0)) * np.float32(0.5)
    q = (0.0, 0.0, float(np.float32(np.sin(np.float64(half)))), float(np.float32(np.cos(np.float64(half)))))
    a = oracle.qrot((0.0, 0.0, 0.0, 1.0), (0.5, 0.0, 0.0)) + np.array([0.0, 0.0, 0.0])
    b = np.float32(oracle.qrot(q, (0.0, -0.5, 0.0))) + np.float32([2.0, 0.0, 0.0])
    assert np.array_equal(a, [0.5, 0.0, 0.0])
    assert np.array_equal(b, np.float32([2.5, -0.0, 0.0]))


# constraints.rs:176-186
def test_constraint_names_cover_all_kinds():
    names = [
        Spring(0, 1, (0, 0, 0), (0, 0, 0), 1.0, 1.0).name,
        Rod(0, 1, (0, 0, 0), (0, 0, 0), 1.0).name,
        BallSocket(0, 1, (0, 0, 0), (0, 0, 0)).name,
        Rope(0, 1, (0, 0, 0), (0, 0, 0), 1.0).name,
    ]
    assert names == ["spring", "rod", "ball and socket", "rope"]


# gpu_physics.rs:702-721
def test_constraint_records_keep_shader_layout_and_kinds():
    cons = [
        Spring(0, 1, (0, 0, 0), (0, 0, 0), 2.0, 1.0),
        Rod(0, 1, (0, 0, 0), (0, 0, 0), 1.0),
        BallSocket(0, 1, (0, 0, 0), (0, 0, 0)),
        Rope(0, 1, (0, 0, 0), (0, 0, 0), 3.0),
    ]
    rec = constraints_from_avbd(cons)
    assert R.AVBD_CONSTRAINT.itemsize == 12 * 4
    assert rec["kind"].tolist() == [R.SPRING, R.ROD, R.BALL_SOCKET, R.ROPE]
    assert rec[3]["anchor_a"][3] == 3.0


# physics/gpu.rs:421-433
@pytest.mark.parametrize("precision", [32, 64])
def test_penetrating_cube_gets_four_floor_contacts(oracle, precision):
    container = AvbdContainer(center=(0.0, 4.0, 0.0), rotation=(0.0, 0.0, 0.0, 1.0), half_extent=4.0)
    body = AvbdBody.cube((0.0, 0.49, 0.0), 0.5, 1.0)
    c = oracle.container_contacts(gpu_bodies([body])[0], state_from_bodies([body])[0], container, AvbdOptions(),
                                  precision=precision)
    assert len(c["a"]) == 4
    assert np.all(c["normal"] == np.array([0.0, 1.0, 0.0]))
    assert np.all(c["r_a"][:, 1] < -0.49)


def _cube_at(x):
    return _body(Shape.cube(0.5), (x, 0.0, 0.0))


# cuboid.rs:404-408
@pytest.mark.parametrize("precision", [32, 64])
def test_touching_cubes_have_zero_penetration(oracle, precision):
    m = oracle.narrow_pair(_cube_at(0.0), _cube_at(1.0), precision=precision)
    assert m and abs(m[0]["depth"]) < 1e-6
    assert m[0]["normal"] @ np.array([1.0, 0.0, 0.0]) > 0.99


# cuboid.rs:410-417
@pytest.mark.parametrize("precision", [32, 64])
def test_overlapping_cubes_report_four_face_contacts(oracle, precision):
    m = oracle.narrow_pair(_cube_at(0.0), _cube_at(0.9), precision=precision)
    assert abs(m[0]["depth"] - 0.1) < 1e-6
    assert m[0]["normal"] @ np.array([1.0, 0.0, 0.0]) > 0.99
    assert len(m) == 4


# cuboid.rs:419-422
@pytest.mark.parametrize("precision", [32, 64])
def test_separated_cubes_have_no_contact(oracle, precision):
    assert oracle.narrow_pair(_cube_at(0.0), _cube_at(1.1), precision=precision) == []

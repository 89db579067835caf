"""Pins of the oracle's instance transforms (the renderer's per-instance
matrices, engine/src/render/renderer.rs:250-255 and context.rs:188-193).

The oracle restates glam's Mat4::from_scale_rotation_translation and Mat4 *
Mat4; glam is not under the reference, so the checks come from what the
mathematics fixes, each through a construction of its own:
- the model matrix's rotation columns are the oracle's glam Quat * Vec3 (a
  different formula: v (w^2 - b.b) + 2 b (v.b) + 2 w (b x v), itself pinned
  by the reference's constraints.rs test) applied to the basis vectors, times
  the scale; the translation column is the position with w = 1;
- the rotation part is orthonormal with determinant +1 for unit quaternions;
- a model matrix maps a local vertex v to position + q (s * v);
- view_proj * model equals the float64 matrix product;
- closed forms: the identity rotation, a rotation about z by theta;
- slot k draws body order[k], bodies repeated or skipped.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import pyoracle


@pytest.fixture(scope="module", autouse=True)
def _load():
    pyoracle.load()


def _quats(rng, n):
    q = rng.normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def _mat(cols16):
    """Column-major 16 values -> 4 x 4 matrix (rows by columns)."""
    return np.asarray(cols16, np.float64).reshape(4, 4).T


def _scene(rng, n):
    pos = rng.uniform(-20, 20, size=(n, 3))
    rot = _quats(rng, n)
    scale = rng.uniform(0.2, 3.0, size=(n, 3))
    vp = rng.normal(size=(4, 4))
    return pos, rot, scale, vp


@pytest.mark.parametrize("precision", [32, 64])
def test_rotation_columns_are_the_rotated_basis(precision):
    rng = np.random.default_rng(1)
    pos, rot, scale, vp = _scene(rng, 64)
    out = pyoracle.instance_transforms(pos, rot, scale, vp.T.reshape(16), precision=precision)
    tol = 2e-6 if precision == 32 else 1e-13
    for k in range(64):
        m = _mat(out[k, 1])
        for axis in range(3):
            e = np.zeros(3)
            e[axis] = 1.0
            want = pyoracle.qrot(rot[k], e, precision=64) * scale[k, axis]
            np.testing.assert_allclose(m[:3, axis], want, atol=tol * (1 + abs(scale[k, axis])))
            assert m[3, axis] == 0.0
        np.testing.assert_allclose(m[:3, 3], pos[k], rtol=0, atol=tol * 20)
        assert m[3, 3] == 1.0


def test_rotation_part_is_orthonormal():
    rng = np.random.default_rng(2)
    pos, rot, _, vp = _scene(rng, 128)
    out = pyoracle.instance_transforms(pos, rot, np.ones((128, 3)), vp.T.reshape(16), precision=64)
    for k in range(128):
        r = _mat(out[k, 1])[:3, :3]
        np.testing.assert_allclose(r.T @ r, np.eye(3), atol=1e-12)
        assert abs(np.linalg.det(r) - 1.0) < 1e-12


def test_model_maps_local_vertices_to_world():
    rng = np.random.default_rng(3)
    pos, rot, scale, vp = _scene(rng, 32)
    out = pyoracle.instance_transforms(pos, rot, scale, vp.T.reshape(16), precision=64)
    for k in range(32):
        m = _mat(out[k, 1])
        for v in rng.normal(size=(4, 3)):
            world = m @ np.append(v, 1.0)
            want = pos[k] + pyoracle.qrot(rot[k], scale[k] * v, precision=64)
            np.testing.assert_allclose(world[:3], want, atol=1e-11)
            assert world[3] == 1.0


@pytest.mark.parametrize("precision", [32, 64])
def test_mvp_is_the_matrix_product(precision):
    rng = np.random.default_rng(4)
    pos, rot, scale, vp = _scene(rng, 64)
    out = pyoracle.instance_transforms(pos, rot, scale, vp.T.reshape(16), precision=precision)
    vp_used = vp.astype(np.float32).astype(np.float64) if precision == 32 else vp
    for k in range(64):
        model = _mat(out[k, 1])
        want = vp_used @ model
        rtol = 2e-6 if precision == 32 else 1e-14
        np.testing.assert_allclose(_mat(out[k, 0]), want, rtol=rtol, atol=rtol * np.abs(want).max())


def test_identity_and_z_rotation_closed_forms():
    th = 0.7
    pos = np.array([[1.0, -2.0, 3.0], [0.5, 0.25, -4.0]])
    rot = np.array([[0.0, 0.0, 0.0, 1.0], [0.0, 0.0, np.sin(th / 2), np.cos(th / 2)]])
    scale = np.array([[2.0, 3.0, 4.0], [1.0, 1.0, 1.0]])
    out = pyoracle.instance_transforms(pos, rot, scale, np.eye(4).reshape(16), precision=64)
    want0 = np.diag([2.0, 3.0, 4.0, 1.0])
    want0[:3, 3] = pos[0]
    np.testing.assert_array_equal(_mat(out[0, 1]), want0)
    np.testing.assert_array_equal(_mat(out[0, 0]), want0)  # identity view_proj
    c, s = np.cos(th), np.sin(th)
    want1 = np.array([[c, -s, 0, 0.5], [s, c, 0, 0.25], [0, 0, 1, -4.0], [0, 0, 0, 1]])
    np.testing.assert_allclose(_mat(out[1, 1]), want1, atol=1e-15)


def test_slots_follow_the_order():
    rng = np.random.default_rng(5)
    pos, rot, scale, vp = _scene(rng, 10)
    order = np.array([3, 3, 0, 9, 5, 3], np.uint32)
    out = pyoracle.instance_transforms(pos, rot, scale, vp.T.reshape(16), order=order)
    each = pyoracle.instance_transforms(pos, rot, scale, vp.T.reshape(16))
    assert out.shape == (6, 2, 4, 4)
    for k, b in enumerate(order):
        np.testing.assert_array_equal(out[k], each[b])

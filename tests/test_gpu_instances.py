"""GPU tests of the instanced-renderer hand-off: dio_world_set_instances /
dio_world_instances (the renderer's per-instance (view_proj * model, model)
pairs, engine/src/render/renderer.rs:222-255 and context.rs:171-199) and
dio_import_external_memory.

The pairs are checked bit for bit against the FP32 oracle (the same float
operations in the same order) from the world's downloaded poses, written
into plain device memory and into memory imported from a DMA-BUF file
descriptor -- the form in which a Vulkan renderer on AMD's drivers exports
its buffer memory (VK_EXTERNAL_MEMORY_HANDLE_TYPE_OPAQUE_FD_BIT).  No Vulkan
exists here; the descriptor comes from hipMemGetHandleForAddressRange of an
ordinary allocation, and the test reads the written bytes back through that
allocation.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

from diomedes_amd import PhysicsContext, scenes
from diomedes_amd._lib import DioError

from .helpers import assert_bits
from diomedes_amd.devmem import DeviceBuffer

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _view_proj(rng) -> np.ndarray:
    """A projection @ view of a camera looking at the field (rows by columns)."""
    f = 1.0 / np.tan(np.radians(60.0) / 2)
    near, far, aspect = 0.1, 500.0, 16.0 / 9.0
    proj = np.array([[f / aspect, 0, 0, 0], [0, -f, 0, 0], [0, 0, far / (near - far), near * far / (near - far)],
                     [0, 0, -1, 0]])
    view = np.eye(4)
    view[:3, 3] = rng.uniform(-40, 40, 3)
    return (proj @ view).astype(np.float32)


def _stepped_field(ctx, seed=3, steps=40):
    scene = scenes.cube_field(seed, nx=16, ny=4, nz=16)
    ctx.upload(scene)
    ctx.step(scene.options, steps)
    return scene


def _expected(oracle, ctx, scene, scale, vp, order):
    st = ctx.download()
    return oracle.instance_transforms(st["positions"], st["orientations"], scale, vp.T.reshape(16), order=order,
                                      precision=32)


def _check(out: np.ndarray, want: np.ndarray, what: str) -> None:
    got = out[: want.size].reshape(want.shape[0], 2, 4, 4)
    assert_bits(got, want, what)


@pytest.mark.parametrize("ordered", [False, True])
def test_instances_match_the_oracle(oracle, ordered):
    _print("instances plain memory, ordered", ordered)
    rng = np.random.default_rng(7)
    with PhysicsContext(1024) as ctx:
        scene = _stepped_field(ctx)
        scale = rng.uniform(0.5, 1.5, size=(scene.n, 3)).astype(np.float32)
        # the renderer's grouped order: a permutation with some bodies drawn twice
        order = np.concatenate([rng.permutation(scene.n), rng.integers(0, scene.n, 100)]).astype(np.uint32) \
            if ordered else None
        count = scene.n if order is None else order.size
        vp = _view_proj(rng)
        ctx.set_instances(scale, order)
        with DeviceBuffer(128 * count + 256) as buf:
            buf.fill(0x7F)
            ctx.instances(vp, buf.ptr, buf.nbytes)
            ctx.synchronize()
            out = buf.read()
            _check(out, _expected(oracle, ctx, scene, scale, vp, order), "instance pairs")
            # nothing past the last slot is written
            assert np.all(out[32 * count:].view(np.uint32) == 0x7F7F7F7F)


def test_instances_through_imported_dma_buf(oracle):
    _print("instances through an imported DMA-BUF")
    rng = np.random.default_rng(8)
    with PhysicsContext(1024) as ctx:
        scene = _stepped_field(ctx, seed=4)
        scale = np.ones((scene.n, 3), np.float32) * np.float32(0.75)
        vp = _view_proj(rng)
        ctx.set_instances(scale)
        want = _expected(oracle, ctx, scene, scale, vp, None)
        nbytes = 128 * scene.n
        with DeviceBuffer(nbytes, exportable=True) as buf:
            buf.fill(0)
            fd = buf.dma_buf_fd()
            try:
                ptr, handle = ctx.import_external(fd, buf.nbytes)
            finally:
                os.close(fd)  # the library holds a duplicate
            assert ptr and ptr != buf.ptr
            ctx.instances(vp, ptr, nbytes)
            ctx.synchronize()
            _check(buf.read(), want, "instance pairs through the import")
            # the mapping follows the world: another step, another frame
            ctx.step(scene.options, 5)
            ctx.instances(vp, ptr, nbytes)
            ctx.release_external(handle)  # waits for the queued writes
            _check(buf.read(), _expected(oracle, ctx, scene, scale, vp, None), "the next frame through the import")
            with pytest.raises(DioError, match="INVALID"):
                ctx.release_external(handle)


def test_import_with_an_offset(oracle):
    _print("import with an offset")
    rng = np.random.default_rng(9)
    with PhysicsContext(1024) as ctx:
        scene = _stepped_field(ctx, seed=5, steps=10)
        scale = rng.uniform(0.5, 2.0, size=(scene.n, 3)).astype(np.float32)
        order = np.arange(0, scene.n, 7, dtype=np.uint32)
        vp = _view_proj(rng)
        ctx.set_instances(scale, order)
        want = _expected(oracle, ctx, scene, scale, vp, order)
        size = 128 * order.size
        offset = 1 << 16
        with DeviceBuffer(offset + size + 4096, exportable=True) as buf:
            buf.fill(0x55)
            fd = buf.dma_buf_fd()
            try:
                ptr, handle = ctx.import_external(fd, buf.nbytes, offset, size)
            finally:
                os.close(fd)
            ctx.instances(vp, ptr, size)
            ctx.synchronize()
            raw = buf.read()
            _check(raw[offset // 4:], want, "instance pairs at the offset")
            assert np.all(raw[: offset // 4].view(np.uint32) == 0x55555555)
            assert np.all(raw[(offset + size) // 4:].view(np.uint32) == 0x55555555)
            ctx.release_external(handle)


def test_hand_off_errors():
    _print("hand-off errors")
    with PhysicsContext(256) as ctx:
        with pytest.raises(DioError, match="NOT_READY"):
            ctx.instances(np.eye(4), 0, 0)
        scene = scenes.box_stack(0)
        ctx.upload(scene)
        scale = np.ones((scene.n, 3), np.float32)
        with pytest.raises(DioError, match="INVALID"):
            ctx.set_instances(scale, np.array([0, scene.n], np.uint32))  # a body past the world
        ctx.set_instances(scale)
        with DeviceBuffer(128 * scene.n + 64) as buf:
            with pytest.raises(DioError, match="CAPACITY") as e:
                ctx.instances(np.eye(4), buf.ptr, 128 * scene.n - 1)
            assert e.value.required == 128 * scene.n
            with pytest.raises(DioError, match="INVALID"):
                ctx.instances(np.eye(4), buf.ptr + 4, 128 * scene.n)  # not 16-byte aligned
            with pytest.raises(DioError, match="INVALID"):
                ctx.instances(np.eye(4), 0, 128 * scene.n)
        with pytest.raises(DioError, match="INVALID"):
            ctx.import_external(-1, 4096)
        with pytest.raises(DioError, match="INVALID"):
            ctx.import_external(0, 4096, 4000, 200)  # past the object
        with pytest.raises(DioError, match="INVALID"):
            ctx.import_external(0, 4096, 0, 0)
        # no slots: nothing to write, any destination
        ctx.set_instances(scale, np.zeros(0, np.uint32))
        ctx.instances(np.eye(4), 0, 0)

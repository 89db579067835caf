"""Raw device memory through the HIP runtime the library links.

For callers of the Python binding that need a device address to hand the
library -- the destination of dio_world_instances -- without another GPU
framework, and to export an allocation as a DMA-BUF file descriptor (the form
a Vulkan renderer on AMD's drivers exports its buffer memory in), which
dio_import_external_memory maps back.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_int, c_size_t, c_uint64, c_void_p

import numpy as np

_hip = None
HIP_MEMCPY_H2D, HIP_MEMCPY_D2H = 1, 2
HIP_MEM_RANGE_HANDLE_TYPE_DMA_BUF_FD = 1


def hip() -> ctypes.CDLL:
    """libamdhip64 by its soname: the instance the library already loaded."""
    global _hip
    if _hip is None:
        from . import _lib

        _lib.load()
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [POINTER(c_void_p), c_size_t]
        h.hipFree.argtypes = [c_void_p]
        h.hipMemcpy.argtypes = [c_void_p, c_void_p, c_size_t, c_int]
        h.hipMemset.argtypes = [c_void_p, c_int, c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        h.hipMemGetHandleForAddressRange.argtypes = [c_void_p, c_void_p, c_size_t, c_int, c_uint64]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipMemset", "hipDeviceSynchronize",
                  "hipMemGetHandleForAddressRange"):
            getattr(h, f).restype = c_int
        _hip = h
    return _hip


def _ok(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


EXPORT_MIN_BYTES = 8 << 20


class DeviceBuffer:
    """A hipMalloc allocation (freed on close).

    exportable: allocate at least EXPORT_MIN_BYTES, in whole 2 MiB units.  The
    runtime places small allocations inside a larger buffer object, and a
    DMA-BUF exported from such a pointer describes the whole object, with the
    pointer at an offset the export does not report; an allocation this large
    is an object of its own, so its DMA-BUF starts at the pointer.
    """

    def __init__(self, nbytes: int, exportable: bool = False):
        if exportable:
            nbytes = max(nbytes, EXPORT_MIN_BYTES)
            nbytes = (nbytes + (2 << 20) - 1) // (2 << 20) * (2 << 20)
        self.nbytes = nbytes
        p = c_void_p()
        _ok(hip().hipMalloc(ctypes.byref(p), nbytes), "hipMalloc")
        self.ptr = p.value

    def fill(self, byte: int) -> None:
        _ok(hip().hipMemset(c_void_p(self.ptr), byte, self.nbytes), "hipMemset")
        _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")

    def read(self, dtype=np.float32, count: int | None = None) -> np.ndarray:
        dt = np.dtype(dtype)
        count = self.nbytes // dt.itemsize if count is None else count
        out = np.empty(count, dt)
        _ok(hip().hipMemcpy(out.ctypes.data_as(c_void_p), c_void_p(self.ptr), out.nbytes, HIP_MEMCPY_D2H),
            "hipMemcpy")
        return out

    def dma_buf_fd(self) -> int:
        """A DMA-BUF file descriptor of the whole allocation (the caller closes it)."""
        fd = c_int(-1)
        _ok(hip().hipMemGetHandleForAddressRange(ctypes.byref(fd), c_void_p(self.ptr), self.nbytes,
                                                  HIP_MEM_RANGE_HANDLE_TYPE_DMA_BUF_FD, 0),
            "hipMemGetHandleForAddressRange")
        return fd.value

    def close(self) -> None:
        if self.ptr:
            hip().hipFree(c_void_p(self.ptr))
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

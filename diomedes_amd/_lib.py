"""ctypes binding of libdiomedes_hip.so (the C ABI in include/diomedes_hip.h).

The library is built in-tree by ``diomedes_amd.build``.  Loading fails loudly
when it is missing: there is no CPU fallback for the physics step.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_uint32, c_uint64, c_void_p
from pathlib import Path

from . import records as R

LIB_PATH = Path(__file__).resolve().parent / "libdiomedes_hip.so"

STATUS = {0: "OK", 1: "CAPACITY", 2: "INVALID", 3: "HIP", 4: "NOT_READY"}

# every exported symbol and its (restype, argtypes)
_VP = c_void_p
_U32P = POINTER(c_uint32)
_FP = POINTER(c_float)
SIGNATURES = {
    "dio_abi_version": (c_uint32, []),
    "dio_default_options": (None, [POINTER(R.Options)]),
    "dio_create": (c_int, [c_int, POINTER(R.Limits), POINTER(_VP)]),
    "dio_destroy": (None, [_VP]),
    "dio_last_error": (c_char_p, [_VP]),
    "dio_required_capacity": (c_uint64, [_VP]),
    "dio_synchronize": (c_int, [_VP]),
    "dio_broad_submit": (c_int, [_VP, _VP, c_uint32]),
    "dio_broad_count": (c_int, [_VP, _U32P]),
    "dio_broad_read_pairs": (c_int, [_VP, _VP, c_uint32, _U32P]),
    "dio_broad_read_aabbs": (c_int, [_VP, _FP, c_uint32]),
    "dio_narrow_submit": (c_int, [_VP, c_uint32]),
    "dio_narrow_submit_pairs": (c_int, [_VP, _VP, c_uint32, _VP, c_uint32]),
    "dio_narrow_read": (c_int, [_VP, _VP, c_uint32, _U32P]),
    "dio_narrow_pairs_noalloc": (c_int, [c_int, _VP, c_uint32, _VP, c_uint32, _VP, c_uint32, _VP, c_uint32]),
    "dio_avbd_submit": (
        c_int,
        [_VP, _VP, _FP, _FP, c_uint32, _VP, c_uint32, _VP, c_uint32, _U32P, _U32P, c_uint32, POINTER(R.RunOptions)],
    ),
    "dio_avbd_read": (c_int, [_VP, _FP, _FP, _FP, _FP, _FP, _FP, _FP]),
    "dio_world_upload": (c_int, [_VP, _VP, _VP, c_uint32, _VP, c_uint32, POINTER(R.Container)]),
    "dio_batch_upload": (
        c_int, [_VP, _VP, _VP, c_uint32, _VP, c_uint32, _U32P, c_uint32, POINTER(R.Container)]
    ),
    "dio_world_upload_hulls": (c_int, [_VP, _FP, c_uint32]),
    "dio_step": (c_int, [_VP, POINTER(R.Options), POINTER(R.StepOpts), c_uint32]),
    "dio_step_get_stats": (c_int, [_VP, POINTER(R.StepStats)]),
    "dio_set_tuning": (c_int, [_VP, POINTER(R.Tuning)]),
    "dio_get_tuning": (c_int, [_VP, POINTER(R.Tuning)]),
    "dio_autotune": (c_int, [_VP, POINTER(R.Options), c_uint32, POINTER(R.Tuning), _FP]),
    "dio_adjoint_begin": (c_int, [_VP, c_uint32, c_uint32, c_uint64]),
    "dio_adjoint_end": (c_int, [_VP]),
    "dio_adjoint_steps": (c_int, [_VP, _U32P]),
    "dio_adjoint_set_mode": (c_int, [_VP, c_uint32]),
    "dio_adjoint_replay": (c_int, [_VP, POINTER(R.Options), _VP, _FP, _FP]),
    "dio_adjoint_backward": (
        c_int, [_VP, POINTER(R.Options), _VP, _FP, _FP, _FP, _FP, _FP, _FP, _FP, _FP, _FP, _FP, _FP, POINTER(c_double)]
    ),
    "dio_world_download": (c_int, [_VP, _FP, _FP, _FP, _FP, _FP]),
    "dio_world_contacts": (c_int, [_VP, _VP, c_uint32, _U32P]),
    "dio_world_pairs": (c_int, [_VP, _VP, c_uint32, _U32P]),
    "dio_world_colours": (c_int, [_VP, _U32P, c_uint32, _U32P]),
    "dio_world_joints": (c_int, [_VP, _FP, _FP, c_uint32, _U32P]),
    "dio_world_set_instances": (c_int, [_VP, _FP, _U32P, c_uint32]),
    "dio_world_instances": (c_int, [_VP, _FP, _VP, c_uint64]),
    "dio_import_external_memory": (c_int, [_VP, c_int, c_uint64, c_uint64, c_uint64, POINTER(_VP), POINTER(_VP)]),
    "dio_release_external_memory": (c_int, [_VP, _VP]),
}

_lib = None


class DioError(RuntimeError):
    """A dio_status other than OK, with the library's message."""

    def __init__(self, status: int, message: str, required: int = 0):
        self.status = STATUS.get(status, str(status))
        self.required = required
        super().__init__(f"{self.status}: {message}")


def load() -> ctypes.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m diomedes_amd.build` (no CPU fallback exists)"
        )
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(ctx, status: int) -> None:
    if status != 0:
        lib = load()
        msg = lib.dio_last_error(ctx).decode(errors="replace") if ctx else "invalid argument"
        req = int(lib.dio_required_capacity(ctx)) if ctx else 0
        raise DioError(status, msg, req)


def ptr(a, ctype=None):
    """ctypes pointer to a contiguous numpy array (None passes through)."""
    if a is None:
        return None
    if ctype is None:
        return a.ctypes.data_as(c_void_p)
    return a.ctypes.data_as(POINTER(ctype))

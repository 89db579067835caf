"""The C ABI: the library loads on the CPU and exports every declared symbol,
the ctypes records match the header's layouts, and argument errors come back
as statuses (no GPU work is issued here)."""

from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from diomedes_amd import _lib
from diomedes_amd import records as R

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "diomedes_hip.h"


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:dio_status|void|uint32_t|uint64_t|const char\*)\s+(dio_\w+)\s*\(", text,
                                 flags=re.M)))


def test_header_declares_the_bound_symbols():
    names = declared_functions()
    assert len(names) >= 20
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_library_loads_and_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (dio_\w+)", out))
    assert set(declared_functions()) <= exported


def _c_sizeof(type_name: str) -> int:
    src = f'#include "{HEADER}"\n#include <stdio.h>\nint main(void){{printf("%zu", sizeof({type_name}));return 0;}}\n'
    tmp = ROOT / "_build"
    tmp.mkdir(exist_ok=True)
    c = tmp / f"sz_{type_name}.c"
    exe = tmp / f"sz_{type_name}"
    c.write_text(src)
    subprocess.run(["gcc", "-o", str(exe), str(c)], check=True)
    return int(subprocess.run([str(exe)], capture_output=True, text=True).stdout)


@pytest.mark.parametrize("ctype,size", [
    ("dio_gpu_body", R.GPU_BODY.itemsize), ("dio_gpu_pair", R.GPU_PAIR.itemsize),
    ("dio_gpu_contact_out", R.GPU_CONTACT_OUT.itemsize), ("dio_avbd_body_state", R.BODY_STATE.itemsize),
    ("dio_avbd_contact", R.AVBD_CONTACT.itemsize), ("dio_avbd_constraint", R.AVBD_CONSTRAINT.itemsize),
    ("dio_run_options", ctypes.sizeof(R.RunOptions)), ("dio_avbd_options", ctypes.sizeof(R.Options)),
    ("dio_container", ctypes.sizeof(R.Container)), ("dio_limits", ctypes.sizeof(R.Limits)),
    ("dio_step_opts", ctypes.sizeof(R.StepOpts)), ("dio_step_stats", ctypes.sizeof(R.StepStats)),
    ("dio_tuning", ctypes.sizeof(R.Tuning)),
])
def test_record_sizes_match_the_header(ctype, size):
    assert _c_sizeof(ctype) == size


def test_reference_record_sizes():
    # Appendix-B sizes of the reference's #[repr(C)] records
    assert (R.GPU_BODY.itemsize, R.GPU_PAIR.itemsize, R.GPU_CONTACT_OUT.itemsize, R.BODY_STATE.itemsize,
            R.AVBD_CONTACT.itemsize, R.AVBD_CONSTRAINT.itemsize) == (112, 8, 80, 80, 144, 48)
    assert R.AVBD_CONTACT.fields["lambda_"][1] == 128 and R.AVBD_CONTACT.fields["c0"][1] == 96
    assert R.GPU_CONTACT_OUT.fields["depth"][1] == 32 and R.GPU_CONTACT_OUT.fields["point_b"][1] == 64


def test_default_options_match_the_reference():
    lib = _lib.load()
    o = R.Options()
    lib.dio_default_options(ctypes.byref(o))
    # physics/gpu.rs:140-155
    assert (o.dt, list(o.gravity), o.iterations) == (np.float32(1 / 60), [0.0, -10.0, 0.0], 10)
    assert (o.alpha, o.beta_lin, o.margin) == (np.float32(0.99), 10000.0, np.float32(0.01))
    assert (o.penalty_min, o.penalty_max, o.gamma, o.friction) == (1.0, np.float32(1e10), np.float32(0.999),
                                                                    np.float32(0.8))
    assert lib.dio_abi_version() == 2


def test_invalid_arguments_return_status_without_a_device():
    lib = _lib.load()
    h = ctypes.c_void_p()
    lim = R.Limits(0, 0, 0, 0, 0)
    assert lib.dio_create(0, ctypes.byref(lim), ctypes.byref(h)) == 2  # max_bodies == 0
    assert lib.dio_create(0, None, ctypes.byref(h)) == 2
    assert lib.dio_synchronize(None) == 2
    assert lib.dio_step(None, None, None, 1) == 2
    assert lib.dio_world_set_instances(None, None, None, 0) == 2
    assert lib.dio_world_instances(None, None, None, 0) == 2
    assert lib.dio_import_external_memory(None, -1, 0, 0, 0, None, None) == 2
    assert lib.dio_release_external_memory(None, None) == 2
    assert lib.dio_last_error(None) == b"null context"

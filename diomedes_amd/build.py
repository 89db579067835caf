"""Build the gfx950 HIP library and the CPU oracle in-tree.

``python -m diomedes_amd.build`` compiles ``diomedes_amd/csrc/*.hip`` into
``diomedes_amd/libdiomedes_hip.so`` with hipcc for gfx950 and, as test
infrastructure, ``oracle/liboracle.so`` with g++ and the plain C caller
``tests/capi/replay_trajectory`` with gcc.

Floating point: every file is built with ``-ffp-contract=off`` and without
fast-math, so each float operation rounds once in source order (the CPU
oracle is built the same way), and with packed FP32 instructions disabled.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libdiomedes_hip.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"
CAPI_DIR = ROOT / "tests" / "capi"
CAPI_EXE = CAPI_DIR / "replay_trajectory"
BUILD_DIR = ROOT / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DIO_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-but-set-variable",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in proc.stdout.splitlines() if "packed-fp32-ops" not in line and line.strip()
    )
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out:
        print(out, file=sys.stderr)


def build_hip(verbose: bool = True) -> Path:
    """Compile every .hip source for gfx950 and link the shared library."""
    BUILD_DIR.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    objs = [BUILD_DIR / (s.stem + ".o") for s in sources]

    def compile_one(pair):
        src, obj = pair
        _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
        return obj

    jobs = min(len(sources), max(1, (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(compile_one, zip(sources, objs)))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB.relative_to(ROOT)}", file=sys.stderr)
    return LIB


def build_oracle(verbose: bool = True) -> Path:
    """Compile the CPU oracle (test infrastructure)."""
    _run(["make", "-C", str(ORACLE_DIR), "-s"])
    if verbose:
        print(f"built {ORACLE_LIB.relative_to(ROOT)}", file=sys.stderr)
    return ORACLE_LIB


def build_capi_caller(verbose: bool = True) -> Path:
    """Compile the plain C99 caller of the C ABI (tests/capi), strict and
    warning-free, linked against the library built beside it."""
    tmp = CAPI_EXE.with_name(CAPI_EXE.name + ".tmp")
    _run(["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", "-pedantic", f"-I{ROOT / 'include'}",
          str(CAPI_DIR / "replay_trajectory.c"), "-o", str(tmp), f"-L{PKG}", "-ldiomedes_hip",
          "-Wl,-rpath,$ORIGIN/../../diomedes_amd"])
    os.replace(tmp, CAPI_EXE)
    if verbose:
        print(f"built {CAPI_EXE.relative_to(ROOT)}", file=sys.stderr)
    return CAPI_EXE


def build_all(verbose: bool = True) -> None:
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    build_hip(verbose)
    build_oracle(verbose)
    build_capi_caller(verbose)


if __name__ == "__main__":
    build_all()

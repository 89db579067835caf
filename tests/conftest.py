"""Shared fixtures: markers, paths, oracle / library availability."""

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle

    pyoracle.load()
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 8)
    pyoracle.threads(int(os.environ.get("ORACLE_THREADS", str(min(16, cores)))))
    return pyoracle


@pytest.fixture(scope="session")
def gpu_ctx_factory():
    from diomedes_amd import PhysicsContext

    def make(n, **kw):
        return PhysicsContext(n, **kw)

    return make

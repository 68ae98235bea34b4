"""Shared test setup.

Markers:
  gpu  - needs a HIP device (MI355X); run with `pytest -m gpu` on a machine with one.
Everything else runs on the CPU: the oracle against closed forms, invariants
and golden fixtures, host logic, and the C ABI library's exports.
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (gfx950)")


def _ensure_built() -> None:
    # Tests load what is built; they build only when nothing is.
    if not (ROOT / "oracle" / "liboracle.so").exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    if not (ROOT / "fullc0ntact_amd" / "libfc_dem.so").exists():
        subprocess.run([sys.executable, "-m", "fullc0ntact_amd.build"], cwd=ROOT, check=True)


_ensure_built()


@pytest.fixture(scope="session")
def gpu_ctx():
    from fullc0ntact_amd import _lib
    if _lib.device_count() == 0:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    ctx = _lib.Context(0)
    yield ctx
    ctx.close()

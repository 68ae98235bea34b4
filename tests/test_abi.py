"""The C ABI library loads and exports every function include/*.h declares;
the host-side logic that needs no GPU behaves (no compute calls here)."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions() -> list[str]:
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(fc_\w+)\s*\(", text, flags=re.M)
        # the FEATFLOW entry points keep the Fortran names
        names += re.findall(r"^\s*void\s+([a-z_0-9]+_)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_stage_entry_points():
    names = declared_functions()
    for required in ("fc_set_parameters", "fc_integrate_system", "fc_calc_hash",
                     "fc_sort_particles", "fc_reorder_data_and_find_cell_start", "fc_collide",
                     "fc_eval_forces",
                     "fc_system_create", "fc_system_update", "fc_system_run",
                     "fc_system_get_array", "fc_system_set_array"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from fullc0ntact_amd import _lib
    missing = [n for n in declared_functions() if not hasattr(_lib.lib, n)]
    assert not missing, missing


def test_binding_covers_declarations():
    from fullc0ntact_amd import _lib
    assert set(_lib.exported_symbols()) >= set(
        n for n in declared_functions() if n.startswith(("fc_system", "fc_context", "fc_")))


def test_sim_params_layout_matches_reference_struct():
    from fullc0ntact_amd import _lib
    P = _lib.SimParams
    assert C.sizeof(P) == 104
    # float3 is 12 bytes unpadded in SimParams (particles_kernel.cuh:27-48)
    assert P.gravity.offset == 16 and P.gridSize.offset == 36 and P.numCells.offset == 48
    assert P.worldOrigin.offset == 52 and P.cellSize.offset == 64
    assert P.spring.offset == 84 and P.boundaryDamping.offset == 100


def test_abi_version_and_error_text():
    from fullc0ntact_amd import _lib
    assert _lib.lib.fc_abi_version() == 1
    assert isinstance(_lib.last_error(), str)


def test_null_arguments_are_rejected_without_a_device():
    from fullc0ntact_amd import _lib
    assert _lib.lib.fc_context_create(0, None) == _lib.FC_ERROR_INVALID_VALUE
    assert _lib.lib.fc_thread_sync(None) == _lib.FC_ERROR_INVALID_VALUE
    assert _lib.lib.fc_system_update(None, C.c_float(0.5)) == _lib.FC_ERROR_INVALID_VALUE
    assert "NULL" in _lib.last_error() or "null" in _lib.last_error()


@pytest.mark.skipif(__import__("fullc0ntact_amd._lib", fromlist=["x"]).device_count() > 0,
                    reason="only meaningful on a host without a GPU")
def test_context_creation_reports_missing_device():
    from fullc0ntact_amd import _lib
    h = C.c_void_p()
    assert _lib.lib.fc_context_create(0, C.byref(h)) == _lib.FC_ERROR_NO_DEVICE
    assert h.value is None


def test_package_imports_lazily():
    import fullc0ntact_amd
    assert "ParticleSystem" in fullc0ntact_amd.__all__
    from fullc0ntact_amd import ParticleSystem  # noqa: F401

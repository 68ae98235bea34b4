"""ctypes binding of libfc_dem.so (the C ABI declared in include/fc_dem.h).

The HIP library is the only compute path: importing this module locates the
in-tree libfc_dem.so and raises if it is missing.  Nothing here falls back to
CPU code.
"""
from __future__ import annotations

import ctypes as C
import weakref
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("FC_DEM_LIB", PKG / "libfc_dem.so"))

FC_SUCCESS = 0
FC_ERROR_INVALID_VALUE = 1
FC_ERROR_GRID_NOT_POW2 = 2
FC_ERROR_HIP = 3
FC_ERROR_OUT_OF_MEMORY = 4
FC_ERROR_NO_DEVICE = 5
FC_ERROR_SORT_STALLED = 6
FC_ERROR_CAPACITY = 7
FC_ERROR_NOT_SUPPORTED = 8

FC_POSITION = 0
FC_VELOCITY = 1
FC_CONFIG_RANDOM = 0
FC_CONFIG_GRID = 1
CELL_EMPTY = 0xFFFFFFFF


class Float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class UInt3(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("z", C.c_uint32)]


class SimParams(C.Structure):
    """fc_sim_params = SimParams (particles_kernel.cuh:27-48), 104 bytes."""
    _fields_ = [
        ("colliderPos", Float3),
        ("colliderRadius", C.c_float),
        ("gravity", Float3),
        ("globalDamping", C.c_float),
        ("particleRadius", C.c_float),
        ("gridSize", UInt3),
        ("numCells", C.c_uint32),
        ("worldOrigin", Float3),
        ("cellSize", Float3),
        ("numBodies", C.c_uint32),
        ("maxParticlesPerCell", C.c_uint32),
        ("spring", C.c_float),
        ("damping", C.c_float),
        ("shear", C.c_float),
        ("attraction", C.c_float),
        ("boundaryDamping", C.c_float),
    ]

    def copy(self) -> "SimParams":
        out = SimParams()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(SimParams))
        return out


assert C.sizeof(SimParams) == 104


class TransInfo(C.Structure):
    """fc_trans_info = TransInfo (inshape3dcore/rigidbody.h:54-66), row-major."""
    _fields_ = [("origin0", Float3), ("origin1", Float3), ("m2w0", C.c_float * 9),
                ("w2m0", C.c_float * 9), ("m2w1", C.c_float * 9)]


class BatchTuning(C.Structure):
    """fc_batch_tuning: launch shape of the tuned batch kernels."""
    _fields_ = [("integratePerThread", C.c_int), ("collideBlock", C.c_int), ("stepMs", C.c_float)]


FC_BOUNDARY_BOX = 0
FC_BOUNDARY_DISTMAP = 1


class StageTiming(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_uint64), ("total_ms", C.c_double)]


class FcError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"fc status {status}: {message}")
        self.status = status


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} not found: build the HIP library first "
            "(python -m fullc0ntact_amd.build or __graft_entry__.build())")
    return C.CDLL(str(LIB_PATH))


lib = _load()

_vp = C.c_void_p
_u32 = C.c_uint32
_st = C.c_int

_SIGS = {
    "fc_context_create": (_st, [C.c_int, C.POINTER(_vp)]),
    "fc_context_destroy": (_st, [_vp]),
    "fc_last_error": (C.c_char_p, []),
    "fc_abi_version": (C.c_int, []),
    "fc_device_count": (C.c_int, []),
    "fc_allocate_array": (_st, [C.POINTER(_vp), C.c_size_t]),
    "fc_free_array": (_st, [_vp]),
    "fc_thread_sync": (_st, [_vp]),
    "fc_copy_array_to_device": (_st, [_vp, _vp, _vp, C.c_size_t, C.c_size_t]),
    "fc_copy_array_from_device": (_st, [_vp, _vp, _vp, C.c_size_t, C.c_size_t]),
    "fc_set_parameters": (_st, [_vp, C.POINTER(SimParams)]),
    "fc_get_parameters": (_st, [_vp, C.POINTER(SimParams)]),
    "fc_integrate_system": (_st, [_vp, _vp, _vp, C.c_float, _u32]),
    "fc_calc_hash": (_st, [_vp, _vp, _vp, _vp, _u32]),
    "fc_sort_particles": (_st, [_vp, _vp, _vp, _u32]),
    "fc_reorder_data_and_find_cell_start": (
        _st, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32]),
    "fc_collide": (_st, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32]),
    "fc_eval_forces": (_st, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32]),
    "fc_sync_probe": (_st, [_vp, _u32, C.c_int, _vp]),
    "fc_sort_particles_noalloc": (_st, [_vp, _vp, _vp, _u32]),
    "fc_sort_particles_noalloc_max": (C.c_size_t, []),
    "fc_system_create": (_st, [_vp, _u32, UInt3, C.POINTER(_vp)]),
    "fc_system_destroy": (_st, [_vp]),
    "fc_system_num_particles": (_u32, [_vp]),
    "fc_system_set_params": (_st, [_vp, C.POINTER(SimParams)]),
    "fc_system_get_params": (_st, [_vp, C.POINTER(SimParams)]),
    "fc_system_set_array": (_st, [_vp, C.c_int, _vp, C.c_int, C.c_int]),
    "fc_system_get_array": (_st, [_vp, C.c_int, _vp]),
    "fc_system_reset": (_st, [_vp, C.c_int, C.c_uint64, C.c_float, C.c_float]),
    "fc_system_update": (_st, [_vp, C.c_float]),
    "fc_system_run": (_st, [_vp, C.c_float, _u32]),
    "fc_system_device_arrays": (_st, [_vp] + [C.POINTER(_vp)] * 8),
    "fc_system_dump_grid": (_st, [_vp, C.POINTER(_u32), C.POINTER(_u32)]),
    "fc_context_enable_timing": (_st, [_vp, C.c_int]),
    "fc_context_reset_timing": (_st, [_vp]),
    "fc_context_get_timing": (_st, [_vp, C.POINTER(StageTiming), C.c_int, C.POINTER(C.c_int)]),
    "fc_copy_probe": (_st, [_vp, _vp, _vp, C.c_size_t, C.c_int, C.POINTER(C.c_float)]),
    "fc_system_set_boundary": (_st, [_vp, C.c_int, _vp]),
    "fc_distmap_create": (_st, [_vp, C.POINTER(C.c_int32), C.c_float, C.POINTER(C.c_float),
                                C.POINTER(C.c_float), _vp, _vp, _vp, _vp, C.POINTER(_vp)]),
    "fc_distmap_build": (_st, [_vp, _vp, _u32, _vp, _u32, C.POINTER(C.c_int32), C.c_float,
                               C.POINTER(C.c_float), C.POINTER(_vp)]),
    "fc_distmap_destroy": (_st, [_vp]),
    "fc_distmap_info": (_st, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_float),
                              C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "fc_distmap_download": (_st, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "fc_distmap_query": (_st, [_vp, _vp, _vp, _u32, _vp, _vp, _vp]),
    "fc_distmap_eval": (_st, [_vp, _vp, _vp, _u32, C.POINTER(TransInfo), _vp, _vp, _vp]),
    "fc_distmap_file_read": (_st, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                   C.POINTER(C.c_int32), _vp, _vp, _vp, _vp, _vp]),
    "fc_distmap_load": (_st, [_vp, C.c_char_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                              C.POINTER(_vp)]),
    "fc_distmap_save": (_st, [_vp, _vp, C.c_char_p]),
    "fc_distmap_get_fbm": (_st, [_vp, _vp]),
    "fc_distmap_set_fbm": (_st, [_vp, _vp]),
    "fc_distmap_query_fbm": (_st, [_vp, _vp, _vp, _u32, _vp]),
    "fc_write_particle_vtk": (_st, [C.c_char_p, _vp, _u32, C.c_float, C.c_int]),
    "fc_mesh_mesh_contacts": (_st, [_vp, _vp, _vp, _u32, C.POINTER(TransInfo), C.c_float, _vp,
                                    _vp, _vp, _u32, C.POINTER(_u32)]),
    "fc_coupling_create": (_st, [_vp, _u32, _vp, UInt3, C.POINTER(_vp)]),
    "fc_coupling_destroy": (_st, [_vp]),
    "fc_coupling_system": (_vp, [_vp]),
    "fc_coupling_set_time_step": (_st, [_vp, C.c_double]),
    "fc_coupling_step": (_st, [_vp, C.c_double, _u32]),
    "fc_coupling_get_body": (_st, [_vp, _u32, _vp, _vp]),
    "fc_coupling_set_body": (_st, [_vp, _u32, _vp, _vp]),
    "fc_featflow_bind": (None, [_vp]),
    "fc_integrate_system_distmap": (_st, [_vp, _vp, _vp, _vp, C.c_float, _u32]),
    "fc_mesh_signed_distance": (_st, [_vp, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _vp]),
    "fc_system_set_polydisperse": (_st, [_vp, C.c_int]),
    "fc_stats_fp64": (_st, [_vp, _vp, _u32, C.POINTER(C.c_double)]),
    "fc_system_stats_fp64": (_st, [_vp, C.POINTER(C.c_double)]),
    "fc_contact_histogram": (_st, [_vp, _vp, _vp, _vp, _u32, C.c_int, _vp, _u32]),
    "fc_system_contact_histogram": (_st, [_vp, _vp, _u32]),
    "fc_batch_create": (_st, [_vp, _u32, _vp, UInt3, C.POINTER(_vp)]),
    "fc_batch_destroy": (_st, [_vp]),
    "fc_batch_num_particles": (_u32, [_vp]),
    "fc_batch_scene_offsets": (_st, [_vp, _vp]),
    "fc_batch_set_params": (_st, [_vp, _vp]),
    "fc_batch_get_params": (_st, [_vp, _vp]),
    "fc_batch_set_array": (_st, [_vp, C.c_int, _vp]),
    "fc_batch_get_array": (_st, [_vp, C.c_int, _vp]),
    "fc_batch_reset": (_st, [_vp, C.c_int, C.c_uint64, C.c_float, C.c_float]),
    "fc_batch_update": (_st, [_vp, C.c_float]),
    "fc_batch_run": (_st, [_vp, C.c_float, _u32]),
    "fc_batch_stats_fp64": (_st, [_vp, _vp]),
    "fc_batch_set_tuning": (_st, [_vp, C.POINTER(BatchTuning)]),
    "fc_batch_get_tuning": (_st, [_vp, C.POINTER(BatchTuning)]),
    "fc_batch_autotune": (_st, [_vp, C.c_float, _u32, C.POINTER(BatchTuning), _vp,
                                C.POINTER(_u32)]),
    "fc_sdf_num_params": (_u32, []),
    "fc_sdf_param_stride": (_u32, []),
    "fc_sdf_create": (_st, [_vp, _u32, _vp, _vp, C.c_float, C.c_uint64, C.POINTER(_vp)]),
    "fc_sdf_destroy": (_st, [_vp]),
    "fc_sdf_get_params": (_st, [_vp, _vp, _vp, _vp]),
    "fc_sdf_set_params": (_st, [_vp, _vp]),
    "fc_sdf_loss_grad": (_st, [_vp, _vp, _vp, _u32, _vp, _vp]),
    "fc_sdf_train": (_st, [_vp, _vp, _vp, _u32, _u32, _u32, C.c_float, C.c_float, _vp]),
    "fc_sdf_map_margin": (_st, [_vp, _u32, _vp, C.c_float, C.c_float, _vp]),
    "fc_sdf_eval": (_st, [_vp, _vp, _vp, _u32, _vp, C.c_int]),
    "fc_mfma_probe": (_st, [_vp, _vp, _vp, _vp]),
    "fc_selftest_pair_math": (_st, [_vp, _vp]),
    "fc_body_set_create": (_st, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, C.c_float, _vp,
                                 C.POINTER(_vp)]),
    "fc_body_set_destroy": (_st, [_vp]),
    "fc_boundary_contacts": (_st, [_vp, _vp, _vp, _vp, _vp, _u32, C.c_float, _vp, _vp, _vp]),
}


def exported_symbols() -> list[str]:
    return list(_SIGS)


for _name, (_res, _args) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


def last_error() -> str:
    msg = lib.fc_last_error()
    return msg.decode() if msg else ""


def check(status: int) -> None:
    if status != FC_SUCCESS:
        raise FcError(status, last_error())


def device_count() -> int:
    return int(lib.fc_device_count())


def ptr(a: np.ndarray) -> int:
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


class Context:
    """fc_context: one HIP stream + one parameter set (setParameters)."""

    def __init__(self, device: int = 0):
        h = _vp()
        check(lib.fc_context_create(device, C.byref(h)))
        self.handle = h
        self._dependents = weakref.WeakSet()

    def register(self, obj) -> None:
        """Objects created on this context (systems, maps, batches,
        surrogates) are closed before the context is destroyed."""
        self._dependents.add(obj)

    def close(self) -> None:
        if self.handle:
            for obj in list(getattr(self, "_dependents", ())):
                try:
                    obj.close()
                except Exception:
                    pass
            check(lib.fc_context_destroy(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self) -> None:
        check(lib.fc_thread_sync(self.handle))

    def set_parameters(self, p: SimParams) -> None:
        check(lib.fc_set_parameters(self.handle, C.byref(p)))

    def get_parameters(self) -> SimParams:
        p = SimParams()
        check(lib.fc_get_parameters(self.handle, C.byref(p)))
        return p

    def sync_probe(self, blocks: int = 256, iterations: int = 1000) -> tuple[float, float]:
        """(us per grid-wide barrier in one cooperative launch, us per kernel
        boundary in a captured graph), for `blocks` blocks of 256 threads."""
        out = (C.c_float * 2)()
        check(lib.fc_sync_probe(self.handle, int(blocks), int(iterations), out))
        return float(out[0]), float(out[1])

    def copy_probe(self, nbytes: int, iterations: int = 20) -> float:
        """Average time (ms) of a plain device-to-device copy of nbytes."""
        src = DeviceArray(self, nbytes)
        dst = DeviceArray(self, nbytes)
        ms = C.c_float(0.0)
        try:
            check(lib.fc_copy_probe(self.handle, dst.ptr, src.ptr, nbytes, iterations, C.byref(ms)))
        finally:
            src.free()
            dst.free()
        return float(ms.value)

    def enable_timing(self, on: bool = True) -> None:
        check(lib.fc_context_enable_timing(self.handle, 1 if on else 0))

    def reset_timing(self) -> None:
        check(lib.fc_context_reset_timing(self.handle))

    def timing(self) -> dict[str, tuple[int, float]]:
        arr = (StageTiming * 64)()
        cnt = C.c_int(0)
        check(lib.fc_context_get_timing(self.handle, arr, 64, C.byref(cnt)))
        return {arr[i].name.decode(): (int(arr[i].launches), float(arr[i].total_ms))
                for i in range(min(cnt.value, 64))}


class DeviceArray:
    """A device buffer allocated through fc_allocate_array."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        h = _vp()
        check(lib.fc_allocate_array(C.byref(h), self.nbytes))
        self.ptr = h.value or 0

    @classmethod
    def from_host(cls, ctx: Context, host: np.ndarray) -> "DeviceArray":
        host = np.ascontiguousarray(host)
        d = cls(ctx, max(host.nbytes, 4))
        if host.nbytes:
            check(lib.fc_copy_array_to_device(ctx.handle, d.ptr, ptr(host), 0, host.nbytes))
        return d

    def upload(self, host: np.ndarray, offset: int = 0) -> None:
        host = np.ascontiguousarray(host)
        check(lib.fc_copy_array_to_device(self.ctx.handle, self.ptr, ptr(host), offset, host.nbytes))

    def download(self, dtype, count: int, offset: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(lib.fc_copy_array_from_device(self.ctx.handle, ptr(out), self.ptr, offset,
                                                out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            check(lib.fc_free_array(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

"""Rigid bodies moved by the particle step (CollisionPipelineGPU,
collision/collisionpipelinegpu.cpp:34-61, with the particle system of
initParticleSystem, fortrancppinterface/cppinterface.cpp:1262-1307), and
the FEATFLOW entry points that read and write the bodies' state."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib


class Coupling:
    """One sphere per rigid body; step() runs `sub_steps` particle updates
    of dt / sub_steps and copies the spheres' state into the bodies."""

    def __init__(self, com, grid=(16, 16, 16), context: _lib.Context | None = None):
        self.ctx = context or _lib.Context()
        c = np.ascontiguousarray(com, np.float64).reshape(-1, 3)
        self.n = len(c)
        h = C.c_void_p()
        check(lib.fc_coupling_create(self.ctx.handle, self.n, c.ctypes.data,
                                     _lib.UInt3(*[int(g) for g in grid]), C.byref(h)))
        self.handle = h
        self.ctx.register(self)

    def close(self) -> None:
        if getattr(self, "handle", None):
            check(lib.fc_coupling_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def system_handle(self):
        return lib.fc_coupling_system(self.handle)

    def set_params(self, params: _lib.SimParams) -> None:
        check(lib.fc_system_set_params(self.system_handle, C.byref(params)))

    def params(self) -> _lib.SimParams:
        p = _lib.SimParams()
        check(lib.fc_system_get_params(self.system_handle, C.byref(p)))
        return p

    def set_time_step(self, dt: float) -> None:
        check(lib.fc_coupling_set_time_step(self.handle, float(dt)))

    def step(self, dt: float, sub_steps: int = 10) -> None:
        check(lib.fc_coupling_step(self.handle, float(dt), int(sub_steps)))

    def body(self, i: int):
        p = np.zeros(3, np.float64)
        v = np.zeros(3, np.float64)
        check(lib.fc_coupling_get_body(self.handle, int(i), p.ctypes.data, v.ctypes.data))
        return p, v

    def set_body(self, i: int, pos=None, vel=None) -> None:
        p = None if pos is None else np.ascontiguousarray(pos, np.float64)
        v = None if vel is None else np.ascontiguousarray(vel, np.float64)
        check(lib.fc_coupling_set_body(self.handle, int(i), p.ctypes.data if p is not None else None,
                                       v.ctypes.data if v is not None else None))

    def bind(self) -> None:
        """Make this the coupling of the FEATFLOW entry points."""
        lib.fc_featflow_bind(self.handle)

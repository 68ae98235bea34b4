"""ParticleSystem: the reference's host particle API over the C ABI.

Mirrors cuda_addon/particleSystem.h:22-136 (method names, argument meaning
and defaults).  All compute runs in libfc_dem.so on the GPU; this class only
keeps the host mirrors m_hPos/m_hVel and the parameter block m_params.

Deliberate differences from the reference (documented in the C ABI header):
  * setArray(POSITION, ...) uploads positions (the reference does nothing
    in non-GL mode, particleSystem.cpp:357-368) and getArray(POSITION)
    downloads them (the reference reads the never-allocated m_dPos,
    particleSystem.cpp:335-339).
  * reset() uses a SplitMix64 stream with an explicit seed instead of
    srand(1973)/rand(), and takes the jitter and velocity amplitude as
    arguments (the reference hard-codes 0.01 r and 0).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import SimParams, check, lib


class ParticleSystem:
    POSITION = _lib.FC_POSITION
    VELOCITY = _lib.FC_VELOCITY
    CONFIG_RANDOM = _lib.FC_CONFIG_RANDOM
    CONFIG_GRID = _lib.FC_CONFIG_GRID

    def __init__(self, numParticles: int, gridSize, bUseOpenGL: bool = False,
                 context: _lib.Context | None = None):
        if bUseOpenGL:
            raise NotImplementedError("OpenGL interop is outside this library")
        self.ctx = context if context is not None else _lib.Context()
        self._own_ctx = context is None
        g = _lib.UInt3(*[int(v) for v in gridSize])
        h = C.c_void_p()
        check(lib.fc_system_create(self.ctx.handle, int(numParticles), g, C.byref(h)))
        self.handle = h
        if context is not None:
            context.register(self)
        self.m_numParticles = int(numParticles)
        self.m_params = SimParams()
        check(lib.fc_system_get_params(self.handle, C.byref(self.m_params)))
        self.m_hPos = np.zeros((self.m_numParticles, 4), np.float32)
        self.m_hVel = np.zeros((self.m_numParticles, 4), np.float32)

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "handle", None):
            check(lib.fc_system_destroy(self.handle))
            self.handle = None
        if getattr(self, "_own_ctx", False) and self.ctx is not None:
            self.ctx.close()
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- stepping ---------------------------------------------------------
    def _push_params(self) -> None:
        # update() re-uploads m_params every step (particleSystem.cpp:236).
        check(lib.fc_system_set_params(self.handle, C.byref(self.m_params)))

    def update(self, deltaTime: float) -> None:
        self._push_params()
        check(lib.fc_system_update(self.handle, C.c_float(deltaTime)))

    def run(self, deltaTime: float, steps: int) -> None:
        """steps consecutive update(deltaTime) calls, graph-launched."""
        self._push_params()
        check(lib.fc_system_run(self.handle, C.c_float(deltaTime), int(steps)))

    def sync(self) -> None:
        self.ctx.sync()

    def setPolydisperse(self, enable: bool = True) -> None:
        """Radii from pos.w (m_params.particleRadius must be the largest
        radius and the cell size at least twice it).  No counterpart in the
        reference, which writes pos.w = 1 and ignores it."""
        check(lib.fc_system_set_polydisperse(self.handle, 1 if enable else 0))

    # -- observables (FP64 reductions on the GPU) -----------------------------
    def stats_fp64(self) -> tuple[float, np.ndarray]:
        """Kinetic energy 1/2 sum |v|^2 and momentum sum v (unit masses)."""
        out = (C.c_double * 4)()
        check(lib.fc_system_stats_fp64(self.handle, out))
        return float(out[0]), np.array(out[1:4], np.float64)

    def contactHistogram(self, nbins: int = 32) -> np.ndarray:
        """Spheres per contact count at the last collide (last bin: >=)."""
        h = np.zeros(nbins, np.uint32)
        check(lib.fc_system_contact_histogram(self.handle, h.ctypes.data, int(nbins)))
        return h

    def setBoundary(self, distance_map=None) -> None:
        """Integrate against the distance-map boundary of
        d_integrateSystem/d_checkBoundaryDistMap (particledem_um.cu:604-638,
        boundaries.cuh:38-89), or against the box walls of integrate_functor
        when distance_map is None.  The map must outlive its use."""
        if distance_map is None:
            check(lib.fc_system_set_boundary(self.handle, _lib.FC_BOUNDARY_BOX, None))
        else:
            check(lib.fc_system_set_boundary(self.handle, _lib.FC_BOUNDARY_DISTMAP,
                                             distance_map.handle))
        self._boundary = distance_map

    # -- initial state ------------------------------------------------------
    def reset(self, config: int, seed: int = 1973, jitter_frac: float = 0.01,
              vel_amp: float = 0.0) -> None:
        self._push_params()
        check(lib.fc_system_reset(self.handle, int(config), int(seed), C.c_float(jitter_frac),
                                  C.c_float(vel_amp)))

    def setParticles(self, positions: np.ndarray, velocities: np.ndarray) -> None:
        """particleSystem.cpp:446-463: w of positions forced to 1, of velocities to 0."""
        p = np.asarray(positions, np.float32).reshape(self.m_numParticles, 4)
        v = np.asarray(velocities, np.float32).reshape(self.m_numParticles, 4)
        self.m_hPos[:, :3] = p[:, :3]
        self.m_hPos[:, 3] = 1.0
        self.m_hVel[:, :3] = v[:, :3]
        self.m_hVel[:, 3] = 0.0
        self.setArray(self.POSITION, self.m_hPos, 0, self.m_numParticles)
        self.setArray(self.VELOCITY, self.m_hVel, 0, self.m_numParticles)

    def setVelocity(self, velocities: np.ndarray) -> None:
        v = np.asarray(velocities, np.float32).reshape(self.m_numParticles, 4)
        self.m_hVel[:, :3] = v[:, :3]
        self.m_hVel[:, 3] = 0.0
        self.setArray(self.VELOCITY, self.m_hVel, 0, self.m_numParticles)

    def getArray(self, array: int) -> np.ndarray:
        host = self.m_hPos if array == self.POSITION else self.m_hVel
        check(lib.fc_system_get_array(self.handle, int(array), host.ctypes.data))
        return host

    def setArray(self, array: int, data: np.ndarray, start: int, count: int) -> None:
        d = np.ascontiguousarray(np.asarray(data, np.float32).reshape(-1, 4)[:count])
        check(lib.fc_system_set_array(self.handle, int(array), d.ctypes.data, int(start),
                                      int(count)))

    def getNumParticles(self) -> int:
        return self.m_numParticles

    def transferArrays(self, start: int = 0, count: int | None = None) -> None:
        """particleSystem.cpp:316-321: refresh the host mirrors."""
        self.getArray(self.POSITION)
        self.getArray(self.VELOCITY)

    def writeParticleVtk(self, path, radius_from_w: bool = False) -> None:
        """The particle dump of particles_gpu (WriteGPUParticleFile,
        vtkwriter.cpp:3399-3446) of the current positions, with the
        system's radius (or pos.w) in the radius column."""
        from .scene_io import write_particle_vtk
        write_particle_vtk(path, self.getArray(self.POSITION), self.getParticleRadius(),
                           radius_from_w)

    def dumpParticles(self, start: int, count: int) -> None:
        self.transferArrays()
        for i in range(start, start + count):
            p, v = self.m_hPos[i], self.m_hVel[i]
            print("pos: (%.4f, %.4f, %.4f, %.4f)" % tuple(p))
            print("vel: (%.4f, %.4f, %.4f, %.4f)" % tuple(v))

    def dumpGrid(self) -> int:
        mx = C.c_uint32(0)
        occ = C.c_uint32(0)
        check(lib.fc_system_dump_grid(self.handle, C.byref(mx), C.byref(occ)))
        print("maximum particles per cell = %d" % mx.value)
        return int(mx.value)

    def occupiedCells(self) -> int:
        mx = C.c_uint32(0)
        occ = C.c_uint32(0)
        check(lib.fc_system_dump_grid(self.handle, C.byref(mx), C.byref(occ)))
        return int(occ.value)

    # -- parameters (particleSystem.h:64-79) --------------------------------
    def setDamping(self, x: float) -> None:
        self.m_params.globalDamping = x

    def setGravity(self, x: float) -> None:
        self.m_params.gravity = _lib.Float3(0.0, 0.0, x)

    def setCollideSpring(self, x: float) -> None:
        self.m_params.spring = x

    def setCollideDamping(self, x: float) -> None:
        self.m_params.damping = x

    def setCollideShear(self, x: float) -> None:
        self.m_params.shear = x

    def setCollideAttraction(self, x: float) -> None:
        self.m_params.attraction = x

    def setColliderPos(self, x) -> None:
        self.m_params.colliderPos = _lib.Float3(*x)

    def getParticleRadius(self) -> float:
        return self.m_params.particleRadius

    def getColliderPos(self):
        c = self.m_params.colliderPos
        return (c.x, c.y, c.z)

    def getColliderRadius(self) -> float:
        return self.m_params.colliderRadius

    def getGridSize(self):
        g = self.m_params.gridSize
        return (g.x, g.y, g.z)

    def getWorldOrigin(self):
        o = self.m_params.worldOrigin
        return (o.x, o.y, o.z)

    def getCellSize(self):
        c = self.m_params.cellSize
        return (c.x, c.y, c.z)

    # -- zero-copy access ---------------------------------------------------
    def device_arrays(self) -> dict[str, int]:
        names = ["pos", "vel", "sortedPos", "sortedVel", "hash", "index", "cellStart", "cellEnd"]
        ptrs = [C.c_void_p() for _ in names]
        check(lib.fc_system_device_arrays(self.handle, *[C.byref(p) for p in ptrs]))
        return {n: (p.value or 0) for n, p in zip(names, ptrs)}

    def download(self, name: str) -> np.ndarray:
        """Copy one internal device array to the host (for checks)."""
        a = self.device_arrays()[name]
        n = self.m_numParticles
        if name in ("pos", "vel", "sortedPos", "sortedVel"):
            out = np.empty((n, 4), np.float32)
        elif name in ("hash", "index"):
            out = np.empty(n, np.uint32)
        else:
            out = np.empty(self.m_params.numCells, np.uint32)
        if out.nbytes:
            check(lib.fc_copy_array_from_device(self.ctx.handle, out.ctypes.data, a, 0, out.nbytes))
        return out


def make_params(n: int, grid: int, radius: float, **overrides) -> SimParams:
    """SimParams with the ParticleSystem defaults for a cubic grid of side
    `grid`, particle radius `radius` and cell = 2r, then field overrides."""
    p = SimParams()
    p.colliderPos = _lib.Float3(-1.2, -0.8, 0.8)
    p.colliderRadius = 0.2
    p.gravity = _lib.Float3(0.0, 0.0, -0.0003)
    p.globalDamping = 1.0
    p.particleRadius = radius
    p.gridSize = _lib.UInt3(grid, grid, grid)
    p.numCells = grid ** 3
    p.worldOrigin = _lib.Float3(-1.0, -1.0, -1.0)
    c = np.float32(radius) * np.float32(2.0)
    p.cellSize = _lib.Float3(c, c, c)
    p.numBodies = n
    p.maxParticlesPerCell = 0
    p.spring = 0.5
    p.damping = 0.02
    p.shear = 0.1
    p.attraction = 0.0
    p.boundaryDamping = -0.5
    for k, v in overrides.items():
        if k in ("colliderPos", "gravity", "worldOrigin", "cellSize"):
            v = _lib.Float3(*v)
        elif k == "gridSize":
            v = _lib.UInt3(*v)
        setattr(p, k, v)
    return p

"""Batch: many independent particle scenes of uneven sizes stepped together
on the GPU, one launch per stage for the whole batch (scene-segmented hash,
graph-launched steady step, autotuned launch shapes).  Each scene's results
are bit-identical to the same scene run alone as a ParticleSystem."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import BatchTuning, SimParams, check, lib


class Batch:
    def __init__(self, scene_counts, gridSize, context: _lib.Context | None = None):
        counts = np.ascontiguousarray(scene_counts, np.uint32)
        self.ctx = context if context is not None else _lib.Context()
        self._own_ctx = context is None
        h = C.c_void_p()
        check(lib.fc_batch_create(self.ctx.handle, len(counts), counts.ctypes.data,
                                  _lib.UInt3(*[int(v) for v in gridSize]), C.byref(h)))
        self.handle = h
        if context is not None:
            context.register(self)
        self.num_scenes = len(counts)
        self.n = int(lib.fc_batch_num_particles(h))
        self.offsets = np.zeros(self.num_scenes + 1, np.uint32)
        check(lib.fc_batch_scene_offsets(h, self.offsets.ctypes.data))

    def close(self) -> None:
        if getattr(self, "handle", None):
            check(lib.fc_batch_destroy(self.handle))
            self.handle = None
        if getattr(self, "_own_ctx", False) and self.ctx is not None:
            self.ctx.close()
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- parameters -----------------------------------------------------------
    def get_params(self) -> list[SimParams]:
        arr = (SimParams * self.num_scenes)()
        check(lib.fc_batch_get_params(self.handle, arr))
        return list(arr)

    def set_params(self, params) -> None:
        arr = (SimParams * self.num_scenes)(*params)
        check(lib.fc_batch_set_params(self.handle, arr))

    # -- state ------------------------------------------------------------------
    def reset(self, config: int = _lib.FC_CONFIG_GRID, seed: int = 1973,
              jitter_frac: float = 0.01, vel_amp: float = 0.0) -> None:
        check(lib.fc_batch_reset(self.handle, int(config), int(seed), C.c_float(jitter_frac),
                                 C.c_float(vel_amp)))

    def set_array(self, array: int, data) -> None:
        d = np.ascontiguousarray(data, np.float32).reshape(self.n, 4)
        check(lib.fc_batch_set_array(self.handle, int(array), d.ctypes.data))

    def get_array(self, array: int) -> np.ndarray:
        out = np.zeros((self.n, 4), np.float32)
        check(lib.fc_batch_get_array(self.handle, int(array), out.ctypes.data))
        return out

    def scene(self, array: np.ndarray, s: int) -> np.ndarray:
        return array[self.offsets[s]:self.offsets[s + 1]]

    # -- stepping -----------------------------------------------------------------
    def update(self, dt: float) -> None:
        check(lib.fc_batch_update(self.handle, C.c_float(dt)))

    def run(self, dt: float, steps: int) -> None:
        check(lib.fc_batch_run(self.handle, C.c_float(dt), int(steps)))

    def sync(self) -> None:
        self.ctx.sync()

    def stats_fp64(self) -> np.ndarray:
        """(num_scenes, 4): kinetic energy, px, py, pz per scene (FP64, GPU)."""
        out = np.zeros((self.num_scenes, 4), np.float64)
        check(lib.fc_batch_stats_fp64(self.handle, out.ctypes.data))
        return out

    # -- launch shapes ---------------------------------------------------------------
    def tuning(self) -> BatchTuning:
        t = BatchTuning()
        check(lib.fc_batch_get_tuning(self.handle, C.byref(t)))
        return t

    def set_tuning(self, integrate_per_thread: int, collide_block: int) -> None:
        t = BatchTuning(integrate_per_thread, collide_block, 0.0)
        check(lib.fc_batch_set_tuning(self.handle, C.byref(t)))

    def autotune(self, dt: float, steps_per_candidate: int = 3):
        """Times the 12 launch shapes on a copy of the state; keeps the best.
        Returns (best tuning, per-candidate ms per step)."""
        best = BatchTuning()
        ms = np.zeros(12, np.float32)
        nc = C.c_uint32(0)
        check(lib.fc_batch_autotune(self.handle, C.c_float(dt), int(steps_per_candidate),
                                    C.byref(best), ms.ctypes.data, C.byref(nc)))
        return best, ms[:nc.value]


def ensemble_scene_counts(num_scenes: int = 256, lo: int = 2000, hi: int = 200000,
                          seed: int = 1973) -> np.ndarray:
    """Scene sizes log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), num_scenes)).astype(np.uint32)


def jittered_params(base: SimParams, num_scenes: int, frac: float = 0.2,
                    seed: int = 1973) -> list[SimParams]:
    """Per-scene copies of `base` with spring and damping scaled by
    U(1-frac, 1+frac)."""
    rng = np.random.default_rng(seed + 1)
    out = []
    for _ in range(num_scenes):
        p = SimParams()
        C.memmove(C.byref(p), C.byref(base), C.sizeof(SimParams))
        p.spring = float(np.float32(base.spring * rng.uniform(1 - frac, 1 + frac)))
        p.damping = float(np.float32(base.damping * rng.uniform(1 - frac, 1 + frac)))
        out.append(p)
    return out

"""The 1M-sphere workload (r = 1/128, 256^3 grid, jittered lattice,
v ~ U(-0.01, 0.01), dt 0.5, seed 1973) run for 1000 graph-launched steps on
the GPU and on the oracle, compared at the end.

The lattice spacing is one diameter, so spheres touch from the first step;
over these steps the lattice falls, is compressed against the floor and
starts to settle, which exercises every branch of the row-range collide
(cut cells, full lists, the direct-loop fallback at x = 0, NaN spheres) at
the size the benchmark measures.  Every position and velocity must equal
the oracle's bit for bit, NaN payloads aside: two spheres clamped into the
same box corner coincide and get a 0 * inf = NaN force on both sides, and
the CPU and the GPU produce different NaN bit patterns for it.
"""
from __future__ import annotations

import os
import time

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

N, GRID, R, DT, STEPS, SEED = 1 << 20, 256, 1.0 / 128, 0.5, 1000, 1973


def same_bits_or_both_nan(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all((a.view(np.uint32) == b.view(np.uint32)) | nan)), int(nan.sum())


def test_million_sphere_run_matches_oracle(gpu_ctx):
    from fullc0ntact_amd import ParticleSystem, make_params
    ps = ParticleSystem(N, (GRID, GRID, GRID), context=gpu_ctx)
    ps.m_params = make_params(N, GRID, R)
    ps.reset(ps.CONFIG_GRID, seed=SEED, jitter_frac=0.1, vel_amp=0.01)
    P = orc.params_from(ps.m_params)
    pos, vel = orc.reset(N, 1, SEED, R, 0.1, 0.01)
    assert np.array_equal(ps.getArray(ps.POSITION).view(np.uint32), pos.view(np.uint32))
    z0 = float(np.mean(pos[:, 2]))
    st = orc.State(pos, vel, P.numCells)

    t0 = time.perf_counter()
    ps.run(DT, STEPS)
    ps.sync()
    t_gpu = time.perf_counter() - t0
    gp, gv = ps.getArray(ps.POSITION), ps.getArray(ps.VELOCITY)
    print(f"GPU: {STEPS} steps of {N} spheres in {t_gpu:.2f} s", flush=True)

    orc.set_threads(max(1, min(16, os.cpu_count() or 1)))
    try:
        t0 = time.perf_counter()
        for done in range(0, STEPS, 250):
            st.run(DT, 250, P)
            print(f"oracle: {done + 250} steps, {time.perf_counter() - t0:.1f} s", flush=True)
    finally:
        orc.set_threads(1)

    ok_p, nan_p = same_bits_or_both_nan(gp, st.pos4)
    ok_v, nan_v = same_bits_or_both_nan(gv, st.vel4)
    print(f"NaN components: positions {nan_p}, velocities {nan_v}", flush=True)
    assert ok_p and ok_v
    # the lattice has fallen
    assert float(np.mean(gp[:, 2][~np.isnan(gp[:, 2])])) < z0
    ps.close()

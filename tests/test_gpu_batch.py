"""Batched independent scenes on the GPU against per-scene oracle runs.

A batch's scene s must equal, bit for bit, the same scene stepped alone
(orc_update with the scene's parameters), whatever the launch shape, through
single updates and through the graph-launched run; a batch of one scene
equals an fc_system.  Per-scene FP64 energies match the oracle's
compensated sums.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GRID = 64


@pytest.fixture(scope="module", autouse=True)
def _threads():
    orc.set_threads(max(1, min(16, os.cpu_count() or 1)))
    yield
    orc.set_threads(1)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _scene_params(counts, seed=5):
    from fullc0ntact_amd import make_params
    from fullc0ntact_amd.batch import jittered_params
    rng = np.random.default_rng(seed)
    base = make_params(0, GRID, 1.0 / 64)
    ps = jittered_params(base, len(counts), 0.2, seed)
    for s, p in enumerate(ps):
        p.numBodies = int(counts[s])
        if s % 3 == 1:  # some scenes with a smaller radius and cell
            r = float(np.float32(1.0 / 96))
            p.particleRadius = r
            c = float(np.float32(r) * np.float32(2.0))
            p.cellSize.x = p.cellSize.y = p.cellSize.z = c
        p.gravity.z = float(np.float32(-0.0003 * rng.uniform(0.5, 2.0)))
    return ps


def _oracle_scenes(batch, params, seed, jitter, vamp, steps, dt):
    out = []
    for s in range(batch.num_scenes):
        n = int(batch.offsets[s + 1] - batch.offsets[s])
        P = orc.params_from(params[s])
        pos, vel = orc.reset(n, 1, seed + s, params[s].particleRadius, jitter, vamp)
        st = orc.State(pos, vel, P.numCells)
        st.run(dt, steps, P)
        out.append(st)
    return out


def _make(gpu_ctx, counts, params, seed=1973, jitter=0.1, vamp=0.01):
    from fullc0ntact_amd.batch import Batch
    b = Batch(counts, (GRID, GRID, GRID), context=gpu_ctx)
    b.set_params(params)
    b.reset(1, seed, jitter, vamp)
    return b


@pytest.mark.parametrize("mode", ["update", "run"])
def test_batch_scenes_match_oracle(gpu_ctx, mode):
    counts = np.array([1, 500, 2003, 4096, 777, 3000, 0, 1500], np.uint32)
    params = _scene_params(counts)
    b = _make(gpu_ctx, counts, params)
    steps, dt = 25, 0.5
    print(f"batch {mode} scenes={len(counts)} n={b.n}", flush=True)
    if mode == "update":
        for _ in range(steps):
            b.update(dt)
    else:
        b.run(dt, steps)
    gp, gv = b.get_array(0), b.get_array(1)
    ref = _oracle_scenes(b, params, 1973, 0.1, 0.01, steps, dt)
    for s, st in enumerate(ref):
        assert np.array_equal(bits(b.scene(gp, s)), bits(st.pos4)), s
        assert np.array_equal(bits(b.scene(gv, s)), bits(st.vel4)), s
    e = b.stats_fp64()
    for s, st in enumerate(ref):
        r = orc.energy_momentum(st.vel4)
        assert abs(e[s, 0] - r[0]) <= 1e-13 * max(r[0], 1e-300), s
        tol = 1e-13 * max(np.abs(st.vel4[:, :3].astype(np.float64)).sum(), 1e-300)
        assert np.all(np.abs(e[s, 1:] - r[1:]) <= tol), s
    b.close()


def test_batch_of_one_scene_equals_system(gpu_ctx):
    from fullc0ntact_amd import ParticleSystem, make_params
    n = 4096
    P = make_params(n, GRID, 1.0 / 64)
    from fullc0ntact_amd.batch import Batch
    b = Batch([n], (GRID, GRID, GRID), context=gpu_ctx)
    b.set_params([P])
    b.reset(1, 1973, 0.1, 0.01)
    b.run(0.5, 20)
    ps = ParticleSystem(n, (GRID, GRID, GRID), context=gpu_ctx)
    ps.m_params = P
    ps.reset(ps.CONFIG_GRID, seed=1973, jitter_frac=0.1, vel_amp=0.01)
    ps.run(0.5, 20)
    assert np.array_equal(bits(b.get_array(0)), bits(ps.getArray(ps.POSITION)))
    assert np.array_equal(bits(b.get_array(1)), bits(ps.getArray(ps.VELOCITY)))
    ps.close()
    b.close()


def test_batch_launch_shapes_do_not_change_results(gpu_ctx):
    counts = np.array([3000, 1200, 5000], np.uint32)
    params = _scene_params(counts, seed=9)
    b = _make(gpu_ctx, counts, params)
    p0, v0 = b.get_array(0), b.get_array(1)
    results = []
    for per in (1, 2, 4, 8):
        for bs in (64, 128, 256):
            b.set_tuning(per, bs)
            b.set_array(0, p0)
            b.set_array(1, v0)
            b.run(0.5, 6)
            results.append((b.get_array(0), b.get_array(1)))
    for p, v in results[1:]:
        assert np.array_equal(bits(p), bits(results[0][0]))
        assert np.array_equal(bits(v), bits(results[0][1]))
    # the autotuner restores the state it measured on
    b.set_array(0, p0)
    b.set_array(1, v0)
    best, ms = b.autotune(0.5, 2)
    assert len(ms) == 12 and np.all(ms > 0)
    assert best.integratePerThread in (1, 2, 4, 8) and best.collideBlock in (64, 128, 256)
    assert abs(best.stepMs - ms.min()) < 1e-6
    assert np.array_equal(bits(b.get_array(0)), bits(p0))
    assert np.array_equal(bits(b.get_array(1)), bits(v0))
    b.run(0.5, 6)
    assert np.array_equal(bits(b.get_array(0)), bits(results[0][0]))
    b.close()


def test_batch_full_size_ensemble_sampled(gpu_ctx):
    """256 scenes, sizes log-uniform in [2000, 200000] (about 11M spheres),
    spring and damping jittered by +-20%: three graph-launched steps, a
    sample of scenes checked against the oracle."""
    from fullc0ntact_amd import make_params
    from fullc0ntact_amd.batch import Batch, ensemble_scene_counts, jittered_params
    counts = ensemble_scene_counts(256, 2000, 200000, seed=1973)
    params = jittered_params(make_params(0, GRID, 1.0 / 64), 256, 0.2, seed=1973)
    for s, p in enumerate(params):
        p.numBodies = int(counts[s])
    b = Batch(counts, (GRID, GRID, GRID), context=gpu_ctx)
    print(f"ensemble n={b.n}", flush=True)
    assert 8e6 < b.n < 1.5e7
    b.set_params(params)
    b.reset(1, 1973, 0.1, 0.01)
    b.run(0.5, 3)
    gp, gv = b.get_array(0), b.get_array(1)
    for s in (0, 1, 77, int(np.argmax(counts)), 255):
        n = int(counts[s])
        P = orc.params_from(params[s])
        pos, vel = orc.reset(n, 1, 1973 + s, params[s].particleRadius, 0.1, 0.01)
        st = orc.State(pos, vel, P.numCells)
        st.run(0.5, 3, P)
        assert np.array_equal(bits(b.scene(gp, s)), bits(st.pos4)), s
        assert np.array_equal(bits(b.scene(gv, s)), bits(st.vel4)), s
    b.close()


@pytest.mark.parametrize("steps", [1, 2, 3, 9, 10, 17])
def test_batch_run_lengths_equal_updates(gpu_ctx, steps):
    """fc_batch_run keeps the state at the sorted slots between its steps,
    fuses each collide with the next integrate, and launches eight steady
    steps per graph; every run length (no steady step, one, a graph and a
    remainder) must equal the same number of single updates, and a run
    after updates and updates after a run must continue the same
    trajectory."""
    counts = np.array([700, 2500, 1, 4096, 0, 1900], np.uint32)
    params = _scene_params(counts, seed=13)
    print(f"[batch run] steps={steps}", flush=True)
    a = _make(gpu_ctx, counts, params)
    b = _make(gpu_ctx, counts, params)
    for _ in range(steps):
        a.update(0.5)
    b.run(0.5, steps)
    assert np.array_equal(bits(a.get_array(0)), bits(b.get_array(0)))
    assert np.array_equal(bits(a.get_array(1)), bits(b.get_array(1)))
    # continue: a run on the updated batch, updates on the run one
    a.run(0.5, 3)
    for _ in range(3):
        b.update(0.5)
    assert np.array_equal(bits(a.get_array(0)), bits(b.get_array(0)))
    assert np.array_equal(bits(a.get_array(1)), bits(b.get_array(1)))
    a.close()
    b.close()

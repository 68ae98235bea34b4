"""Long runs of the 4096-sphere case (r = 1/64, 64^3 grid, jittered
lattice, dt 0.5, 1000 steps) on the GPU against the oracle, for several
seeds.

DEM trajectories are chaotic, so a method that is only close to the
reference drifts apart from it within a few hundred steps and can only be
compared through ensemble statistics.  This build is bit-exact, so the
trajectories do not drift at all: after 1000 graph-launched steps every
position and velocity equals the oracle's, and so do the ensemble
statistics of the settled bed -- the kinetic-energy decay curve (FP64,
within the reduction's 1e-13), the packing fraction and the contact-count
histogram.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from fullc0ntact_amd import statistics
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

N, GRID, R, DT, STEPS, EVERY = 4096, 64, 1.0 / 64, 0.5, 1000, 100


@pytest.fixture(scope="module", autouse=True)
def _threads():
    orc.set_threads(max(1, min(16, os.cpu_count() or 1)))
    yield
    orc.set_threads(1)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("seed", [1973, 7, 4242])
def test_long_run_is_bit_exact_with_equal_statistics(gpu_ctx, seed):
    from fullc0ntact_amd import ParticleSystem, make_params
    ps = ParticleSystem(N, (GRID, GRID, GRID), context=gpu_ctx)
    ps.m_params = make_params(N, GRID, R)
    ps.reset(ps.CONFIG_GRID, seed=seed, jitter_frac=0.1, vel_amp=0.01)
    P = orc.params_from(ps.m_params)
    pos, vel = orc.reset(N, 1, seed, R, 0.1, 0.01)
    st = orc.State(pos, vel, P.numCells)

    e_gpu = statistics.energy_curve(ps, DT, STEPS, EVERY)
    e_ref = []
    for _ in range(STEPS // EVERY):
        st.run(DT, EVERY, P)
        e_ref.append(orc.energy_momentum(st.vel4)[0])
    e_ref = np.array(e_ref)
    print(f"seed {seed}: energy {e_ref[0]:.4e} -> {e_ref[-1]:.4e}", flush=True)

    gp, gv = ps.getArray(ps.POSITION), ps.getArray(ps.VELOCITY)
    assert np.array_equal(bits(gp), bits(st.pos4))
    assert np.array_equal(bits(gv), bits(st.vel4))
    assert np.all(np.abs(e_gpu - e_ref) <= 1e-13 * np.abs(e_ref))

    phi_gpu = statistics.packing_fraction(gp, radius=R)
    phi_ref = statistics.packing_fraction(st.pos4, radius=R)
    assert phi_gpu == phi_ref and 0.05 < phi_ref < 0.75

    # histogram of the last collide: the oracle's sorted state of the same
    # step is the one its last update left
    h = ps.contactHistogram(16)
    href = orc.contact_histogram(st.sortedPos4, st.cellStart, st.cellEnd, P, 16)
    assert np.array_equal(h, href) and h.sum() == N
    ps.close()

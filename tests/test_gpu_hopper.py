"""The polydisperse hopper at full size: 4,194,304 spheres with log-normal
radii (median 1/512, sigma 0.25, clipped to a 3x ratio), cell = the largest
diameter, a 512^3 grid (134M cells, the next power of two covering the
box), and a (2, 3) torus-knot tube of 1024 x 32 x 2 = 65,536 triangles as
the rigid boundary through an FP32 256^3 distance map over [-1, 1]^3.

The spheres start on a lattice that falls onto the knot; some start inside
the tube and are pushed out by the boundary.  A few graph-launched steps on
the GPU must equal the oracle's polydisperse map-boundary steps bit for bit
at this size (the same kernels are pinned against the oracle on small
cases; this checks the 512^3 tables, the non-power-of-two cell size and the
4M-sphere sort).  The oracle's map is the GPU-built one, downloaded; the
map builder itself is pinned on small meshes.
"""
from __future__ import annotations

import os
import time

import numpy as np
import pytest

from fullc0ntact_amd import meshes, statistics
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

N, GRID, MEDIAN, STEPS = 1 << 22, 512, 1.0 / 512, 6


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_hopper_4m_polydisperse_torus_knot_matches_oracle(gpu_ctx):
    from fullc0ntact_amd import ParticleSystem, make_params
    from fullc0ntact_amd.distance_map import DistanceMap
    rmax = float(np.float32(MEDIAN * np.sqrt(3.0)))
    radii = statistics.polydisperse_radii(N, MEDIAN, 0.25, 3.0, seed=5)
    assert float(radii.max()) <= rmax
    pos, vel = orc.reset(N, 1, 1973, rmax, 0.1, 0.01)
    pos[:, 3] = radii

    t0 = time.perf_counter()
    v, f = meshes.torus_knot(2, 3, 1024, 32, radius=0.12, scale=0.25, center=(0.0, 0.0, -0.55))
    assert len(f) == 65536
    gm = DistanceMap.from_mesh(gpu_ctx, v, f, (256, 256, 256), 2.0 / 256, (-1.0, -1.0, -1.0))
    lay = gm.download()
    m = orc.DistMap.from_reference_layout(gm.cells, np.float32(gm.cell_size),
                                          np.asarray(gm.bmin, np.float32),
                                          np.asarray(gm.bmax, np.float32), lay["vertex_coords"],
                                          lay["normals"], lay["contact_points"], lay["distance"])
    del lay
    print(f"map: {len(f)} triangles, 257^3 vertices, {time.perf_counter() - t0:.1f} s",
          flush=True)

    ps = ParticleSystem(N, (GRID, GRID, GRID), context=gpu_ctx)
    ps.m_params = make_params(N, GRID, rmax)
    ps.setGravity(-0.003)
    ps.setArray(ps.POSITION, pos, 0, N)
    ps.setArray(ps.VELOCITY, vel, 0, N)
    ps.setPolydisperse(True)
    ps.setBoundary(gm)
    P = orc.params_from(ps.m_params)
    assert P.numCells == GRID ** 3
    st = orc.State(pos, vel, P.numCells)

    t0 = time.perf_counter()
    ps.run(0.5, STEPS)
    gp, gv = ps.getArray(ps.POSITION), ps.getArray(ps.VELOCITY)
    print(f"GPU: {STEPS} steps in {time.perf_counter() - t0:.2f} s", flush=True)

    orc.set_threads(max(1, min(16, os.cpu_count() or 1)))
    try:
        t0 = time.perf_counter()
        for i in range(STEPS):
            orc.update_ex(st, 0.5, P, m, poly=True)
        print(f"oracle: {STEPS} steps in {time.perf_counter() - t0:.1f} s", flush=True)
    finally:
        orc.set_threads(1)
    assert np.array_equal(bits(gp), bits(st.pos4))
    assert np.array_equal(bits(gv), bits(st.vel4))
    # the knot pushed spheres out and they collided: velocities changed
    assert np.any(gv[:, :3] != vel[:, :3])
    ps.setBoundary(None)
    ps.close()
    gm.close()

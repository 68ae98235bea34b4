"""GPU observables and polydisperse spheres against the oracle.

 - FP64 energy/momentum reduction: within 1e-13 relative of the oracle's
   compensated sums, and bitwise repeatable;
 - contact-count histograms: exact integers, equal to the oracle's;
 - polydisperse steps (radius in pos.w): with equal radii bit-identical to
   the monodisperse step; with log-normal radii bit-identical to the
   oracle, through update() and the graph-launched run(), with the box and
   with the distance-map boundary; a radius above particleRadius is
   reported as an error.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import distmap_cases as dc
from fullc0ntact_amd import meshes, statistics
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _threads():
    orc.set_threads(max(1, min(16, os.cpu_count() or 1)))
    yield
    orc.set_threads(1)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _system(gpu_ctx, n, grid, r, **kw):
    from fullc0ntact_amd import ParticleSystem, make_params
    ps = ParticleSystem(n, (grid, grid, grid), context=gpu_ctx)
    ps.m_params = make_params(n, grid, r, **kw)
    return ps


def test_energy_momentum_fp64(gpu_ctx):
    from fullc0ntact_amd import _lib
    rng = np.random.default_rng(5)
    for n in (1, 255, 4096, 1000003):
        print(f"stats n={n}", flush=True)
        vel = np.zeros((n, 4), np.float32)
        vel[:, :3] = rng.normal(0.0, 0.01, size=(n, 3))
        vel[:, 3] = 7.0  # ignored
        dv = _lib.DeviceArray.from_host(gpu_ctx, vel)
        out = (_lib.C.c_double * 4)()
        _lib.check(_lib.lib.fc_stats_fp64(gpu_ctx.handle, dv.ptr, n, out))
        g = np.array(out[:])
        _lib.check(_lib.lib.fc_stats_fp64(gpu_ctx.handle, dv.ptr, n, out))
        assert np.array_equal(np.array(out[:]).view(np.uint64), g.view(np.uint64))
        ref = orc.energy_momentum(vel)
        scale = np.array([ref[0], *np.abs(vel[:, :3].astype(np.float64)).sum(0)])
        assert np.all(np.abs(g - ref) <= 1e-13 * scale), (g, ref)


def test_system_stats_and_contact_histogram(gpu_ctx):
    n, grid, r = 4096, 64, 1.0 / 64
    ps = _system(gpu_ctx, n, grid, r)
    ps.reset(ps.CONFIG_GRID, seed=1973, jitter_frac=0.1, vel_amp=0.01)
    P = orc.params_from(ps.m_params)
    pos, vel = orc.reset(n, 1, 1973, r, 0.1, 0.01)
    st = orc.State(pos, vel, P.numCells)
    for _ in range(30):
        ps.update(0.5)
        st.update(0.5, P)
    e, p = ps.stats_fp64()
    ref = orc.energy_momentum(st.vel4)
    assert abs(e - ref[0]) <= 1e-13 * ref[0]
    assert np.allclose(p, ref[1:], rtol=0, atol=1e-13 * np.abs(st.vel4[:, :3]).sum())
    h = ps.contactHistogram(16)
    href = orc.contact_histogram(st.sortedPos4, st.cellStart, st.cellEnd, P, 16)
    assert np.array_equal(h, href) and h.sum() == n and h[1:].sum() > 0
    ps.close()


def test_contact_histogram_stage(gpu_ctx):
    from fullc0ntact_amd import _lib
    n, r = 5003, 1.0 / 64
    P = orc.make_params(n, 64, r)
    from fullc0ntact_amd import make_params
    gpu_ctx.set_parameters(make_params(n, 64, r))
    rng = np.random.default_rng(2)
    pos = np.ones((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-0.5, 0.5, size=(n, 3))
    h, i = orc.calc_hash(pos, P)
    h, i = orc.sort_pairs(h, i)
    cs, ce, sp, _ = orc.reorder(h, i, pos, pos, P.numCells)
    d = [_lib.DeviceArray.from_host(gpu_ctx, x) for x in (sp, cs, ce)]
    dh = _lib.DeviceArray(gpu_ctx, 4 * 40)
    _lib.check(_lib.lib.fc_contact_histogram(gpu_ctx.handle, d[0].ptr, d[1].ptr, d[2].ptr, n, 0,
                                             dh.ptr, 40))
    gpu_ctx.sync()
    got = dh.download(np.uint32, 40)
    assert np.array_equal(got, orc.contact_histogram(sp, cs, ce, P, 40))


def test_polydisperse_equal_radii_is_monodisperse(gpu_ctx):
    n, grid, r = 4096, 64, 1.0 / 64
    pos, vel = orc.reset(n, 1, 1973, r, 0.1, 0.01)
    pos[:, 3] = np.float32(r)
    out = []
    for poly in (False, True):
        ps = _system(gpu_ctx, n, grid, r)
        ps.setArray(ps.POSITION, pos, 0, n)
        ps.setArray(ps.VELOCITY, vel, 0, n)
        ps.setPolydisperse(poly)
        ps.run(0.5, 25)
        out.append((ps.getArray(ps.POSITION).copy(), ps.getArray(ps.VELOCITY).copy()))
        ps.close()
    assert np.array_equal(bits(out[0][0]), bits(out[1][0]))
    assert np.array_equal(bits(out[0][1]), bits(out[1][1]))


def _poly_state(n, rmax, seed=11):
    radii = statistics.polydisperse_radii(n, rmax / np.sqrt(3.0), 0.25, 3.0, seed=seed)
    radii = np.minimum(radii, np.float32(rmax))
    pos, vel = orc.reset(n, 1, 1973, float(rmax), 0.1, 0.01)
    pos[:, 3] = radii
    return pos, vel


@pytest.mark.parametrize("n", [8000, 24000])  # row-split and row-range collides
@pytest.mark.parametrize("mode", ["update", "run"])
def test_polydisperse_matches_oracle(gpu_ctx, mode, n):
    grid, rmax = 64, 1.0 / 64
    pos, vel = _poly_state(n, rmax)
    ps = _system(gpu_ctx, n, grid, rmax)
    ps.setArray(ps.POSITION, pos, 0, n)
    ps.setArray(ps.VELOCITY, vel, 0, n)
    ps.setPolydisperse(True)
    P = orc.params_from(ps.m_params)
    st = orc.State(pos, vel, P.numCells)
    steps = 40
    if mode == "update":
        for _ in range(steps):
            ps.update(0.5)
    else:
        ps.run(0.5, steps)
    for _ in range(steps):
        orc.update_ex(st, 0.5, P, None, poly=True)
    assert np.array_equal(bits(ps.getArray(ps.POSITION)), bits(st.pos4))
    assert np.array_equal(bits(ps.getArray(ps.VELOCITY)), bits(st.vel4))
    h = ps.contactHistogram(24)
    assert np.array_equal(h, orc.contact_histogram(st.sortedPos4, st.cellStart, st.cellEnd, P,
                                                   24, poly=True))
    ps.close()


def test_polydisperse_distmap_boundary_matches_oracle(gpu_ctx):
    from fullc0ntact_amd.distance_map import DistanceMap
    n, grid, rmax = 6000, 64, 1.0 / 64
    pos, vel = _poly_state(n, rmax, seed=3)
    pos[:, 2] += 0.8  # start above the boundary body
    v, f = meshes.torus_knot(2, 3, 256, 16, radius=0.15, scale=0.25, center=(0.0, 0.0, -0.4))
    gm = DistanceMap.from_mesh(gpu_ctx, v, f, (48, 48, 48), 2.0 / 48, (-1.0, -1.0, -1.0))
    lay = gm.download()
    m = orc.DistMap.from_reference_layout(gm.cells, np.float32(gm.cell_size),
                                          np.asarray(gm.bmin, np.float32),
                                          np.asarray(gm.bmax, np.float32), lay["vertex_coords"],
                                          lay["normals"], lay["contact_points"], lay["distance"])
    ps = _system(gpu_ctx, n, grid, rmax)
    ps.setGravity(-0.003)
    ps.setArray(ps.POSITION, pos, 0, n)
    ps.setArray(ps.VELOCITY, vel, 0, n)
    ps.setPolydisperse(True)
    ps.setBoundary(gm)
    P = orc.params_from(ps.m_params)
    st = orc.State(pos, vel, P.numCells)
    steps = 40
    ps.run(0.5, steps)
    for _ in range(steps):
        orc.update_ex(st, 0.5, P, m, poly=True)
    assert np.array_equal(bits(ps.getArray(ps.POSITION)), bits(st.pos4))
    assert np.array_equal(bits(ps.getArray(ps.VELOCITY)), bits(st.vel4))
    ps.setBoundary(None)
    ps.close()
    gm.close()


def test_polydisperse_radius_above_bound_is_reported(gpu_ctx):
    from fullc0ntact_amd import _lib
    n, grid, r = 512, 64, 1.0 / 64
    pos, vel = orc.reset(n, 1, 1973, r, 0.1, 0.0)
    pos[:, 3] = np.float32(r)
    pos[7, 3] = np.float32(1.5 * r)
    ps = _system(gpu_ctx, n, grid, r)
    ps.setArray(ps.POSITION, pos, 0, n)
    ps.setArray(ps.VELOCITY, vel, 0, n)
    ps.setPolydisperse(True)
    ps.update(0.5)
    with pytest.raises(_lib.FcError):
        ps.getArray(ps.POSITION)
    # the error word is cleared once reported
    ps.getArray(ps.POSITION)
    ps.close()

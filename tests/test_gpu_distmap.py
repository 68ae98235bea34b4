"""Parity of the distance-map path on the GPU (through the C ABI) with the
oracle.

 - the map builder (exact point-mesh distance, binary64, stored in binary32)
   and the free-point signed distance: bit for bit;
 - queryMap, eval_distmap and the distance-map boundary of integrateSystem on
   the built map: bit for bit (the GPU keeps the reference's mixed
   double/float arithmetic and is compiled without contraction);
 - whole steps of a particle system with the distance-map boundary, through
   update() and the graph-launched run(): bit for bit;
 - the reference's own GPU checks, test_dist_comp (cuda_addon/
   unit_tests.cuh:290-306) and test_inside (difi.cu:1252-1270).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

import distmap_cases as dc
from fullc0ntact_amd import meshes
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _oracle_threads():
    orc.set_threads(max(1, min(16, os.cpu_count() or 1)))
    yield
    orc.set_threads(1)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _build(ctx, verts, tris, cells, cs, bmin):
    from fullc0ntact_amd.distance_map import DistanceMap
    print(f"build map cells={cells} tris={len(tris)}", flush=True)
    return DistanceMap.from_mesh(ctx, verts, tris, cells, cs, bmin)


def _oracle_from(gpu_map):
    lay = gpu_map.download()
    m = orc.DistMap.from_reference_layout(gpu_map.cells, np.float32(gpu_map.cell_size),
                                          np.asarray(gpu_map.bmin, np.float32),
                                          np.asarray(gpu_map.bmax, np.float32),
                                          lay["vertex_coords"], lay["normals"],
                                          lay["contact_points"], lay["distance"])
    return m, lay


MESHES = {
    "sphere": (lambda: dc.small_sphere(0.5, 24, 12), (24, 24, 24), 1.5 / 24, (-0.75,) * 3),
    "hopper": (lambda: dc.golden_mesh("hopper"), (28, 28, 20), 0.075, (-1.05, -1.05, -0.55)),
    "torusknot1": (lambda: dc.golden_mesh("torusknot1"), (20, 16, 20), 0.8,
                   (-8.0, -6.0, -8.0)),
}


@pytest.mark.parametrize("name", list(MESHES))
def test_builder_bitexact(gpu_ctx, name):
    make, cells, cs, bmin = MESHES[name]
    v, f = make()
    gm = _build(gpu_ctx, v, f, cells, cs, bmin)
    lay = gm.download()
    _, ref, bmax = dc.oracle_map(v, f, cells, cs, bmin)
    assert np.array_equal(np.asarray(gm.bmax, np.float32), bmax)
    for k in ("vertex_coords", "contact_points", "distance", "normals"):
        assert np.array_equal(bits(lay[k]), bits(ref[k])), k
    # test_inside (difi.cu:1252-1270): inside counts agree
    assert (lay["distance"] < 0).sum() == (ref["distance"] < 0).sum() > 0
    gm.close()


def test_builder_full_size_sampled(gpu_ctx):
    """A (2,3) torus-knot tube of 1024 x 32 x 2 = 65 536 triangles on a 64^3
    lattice; a sample of lattice vertices is checked against the oracle."""
    v, f = meshes.torus_knot(2, 3, 1024, 32, radius=0.2, scale=0.3)
    cells, cs, bmin = (64, 64, 64), 2.0 / 64, (-1.0, -1.0, -1.0)
    gm = _build(gpu_ctx, v, f, cells, cs, bmin)
    lay = gm.download()
    q = dc.lattice_points(cells, cs, bmin)
    rng = np.random.default_rng(5)
    sel = rng.choice(len(q), 600, replace=False)
    sdf, cp, _ = orc.mesh_sdf(v, f, q[sel])
    assert np.array_equal(bits(lay["distance"][sel]), bits(sdf.astype(np.float32)))
    assert np.array_equal(bits(lay["contact_points"][sel]), bits(cp.astype(np.float32)))
    inside = (lay["distance"] < 0).mean()
    # the tube volume over the box volume, loosely
    vol = meshes.signed_volume(v, f)
    assert abs(inside - vol / 8.0) < 0.02
    gm.close()


def test_mesh_signed_distance_bitexact(gpu_ctx):
    from fullc0ntact_amd.distance_map import mesh_signed_distance
    v, f = dc.golden_mesh("hopper")
    rng = np.random.default_rng(11)
    p = rng.uniform(v.min(0) - 0.3, v.max(0) + 0.3, size=(20000, 3))
    sd, cp = mesh_signed_distance(gpu_ctx, v, f, p)
    rsd, rcp, _ = orc.mesh_sdf(v, f, p)
    assert np.array_equal(sd.view(np.uint64), rsd.view(np.uint64))
    assert np.array_equal(cp.view(np.uint64), rcp.view(np.uint64))


@pytest.fixture(scope="module")
def sphere_map(gpu_ctx):
    v, f = dc.small_sphere(0.5, 24, 12)
    gm = _build(gpu_ctx, v, f, (32, 32, 32), 1.5 / 32, (-0.75,) * 3)
    m, lay = _oracle_from(gm)
    yield gm, m, lay, (v, f)
    gm.close()


def _query_points(rng, n=50000):
    pts = rng.uniform(-0.9, 0.9, size=(n, 3)).astype(np.float32)
    # faces, edges and corners of the box, vertices of the lattice
    pts[:6, :] = [[0.75, 0.1, 0.2], [-0.75, 0.1, 0.2], [0.1, 0.75, 0.2], [0.1, -0.75, 0.2],
                  [0.1, 0.2, 0.75], [0.75, 0.75, 0.75]]
    pts[6] = [0.001, 0.0, 0.0]
    pts[7] = [np.nan, 0.0, 0.0]
    return pts


def test_query_bitexact(gpu_ctx, sphere_map):
    gm, m, lay, (v, f) = sphere_map
    pts = _query_points(np.random.default_rng(3))
    d, cp, nr = gm.query(pts)
    rd, rcp, rnr = orc.distmap_query(m, pts)
    assert np.array_equal(bits(d), bits(rd))
    assert np.array_equal(bits(cp), bits(rcp))
    assert np.array_equal(bits(nr), bits(rnr))
    assert np.all(d[np.any(np.abs(pts) > 0.75, axis=1)] == 5.0)
    # test_dist_comp (unit_tests.cuh:290-306): GPU query at (0.001, 0, 0)
    # against the CPU value, 1e-3 relative (here: identical)
    assert abs((d[6] - rd[6]) / rd[6]) < 1e-3
    # test_inside: GPU and CPU inside counts over the lattice vertices
    dv, _, _ = gm.query(lay["vertex_coords"])
    assert (dv < 0).sum() == (lay["distance"] < 0).sum()


def _rot(ax, ang):
    c, s = np.cos(ang), np.sin(ang)
    if ax == "z":
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def test_eval_bitexact(gpu_ctx, sphere_map):
    from fullc0ntact_amd import _lib
    gm, m, lay, _ = sphere_map
    R0 = (_rot("z", 0.4) @ _rot("x", 0.1)).astype(np.float32)
    R1 = (_rot("x", -0.3)).astype(np.float32)
    o0, o1 = (0.05, -0.02, 0.1), (0.0, 0.1, -0.05)
    gi, oi = _lib.TransInfo(), orc.OrcTransInfo()
    gi.origin0, gi.origin1 = _lib.Float3(*o0), _lib.Float3(*o1)
    for a in range(3):
        oi.origin0[a], oi.origin1[a] = o0[a], o1[a]
    for t in (gi, oi):
        for k in range(9):
            t.m2w0[k] = float(R0.ravel()[k])
            t.w2m0[k] = float(R0.T.ravel()[k])
            t.m2w1[k] = float(R1.ravel()[k])
    pts = _query_points(np.random.default_rng(4), 20000)
    cps, nr, d = gm.eval(pts, gi)
    rcps, rnr, rd = orc.distmap_eval(m, pts, oi)
    assert np.array_equal(bits(d), bits(rd))
    assert np.array_equal(bits(cps), bits(rcps))
    assert np.array_equal(bits(nr), bits(rnr))
    assert (d == 1000.0).sum() > 0 and (d < 1000.0).sum() > 0


def test_integrate_distmap_bitexact(gpu_ctx, sphere_map):
    from fullc0ntact_amd import _lib, make_params
    gm, m, lay, _ = sphere_map
    n = 30000
    P = make_params(n, 64, 1.0 / 64)
    gpu_ctx.set_parameters(P)
    rng = np.random.default_rng(8)
    pos = np.ones((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-1.1, 1.1, size=(n, 3))
    pos[:100, 2] = rng.uniform(9.9, 10.1, size=100)
    pos[100:200, 2] = rng.uniform(-1.1, -0.9, size=100)
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.uniform(-0.01, 0.01, size=(n, 3))
    vel[:, 3] = rng.uniform(0, 1, size=n)
    dp = _lib.DeviceArray.from_host(gpu_ctx, pos)
    dv = _lib.DeviceArray.from_host(gpu_ctx, vel)
    _lib.check(_lib.lib.fc_integrate_system_distmap(gpu_ctx.handle, gm.handle, dp.ptr, dv.ptr,
                                                    C.c_float(0.5), n))
    gpu_ctx.sync()
    gp = dp.download(np.float32, 4 * n).reshape(n, 4)
    gv = dv.download(np.float32, 4 * n).reshape(n, 4)
    rp, rv = orc.integrate_distmap(pos, vel, 0.5, orc.params_from(P), m)
    assert np.array_equal(bits(gp), bits(rp))
    assert np.array_equal(bits(gv), bits(rv))
    moved = np.any(rv[:, :3] != vel[:, :3], axis=1)
    assert moved.sum() > 1000  # the boundaries were exercised


def _knot_scene(n, r):
    """Particles on a jittered lattice above a torus-knot tube boundary."""
    v, f = meshes.torus_knot(2, 3, 256, 16, radius=0.15, scale=0.25, center=(0.0, 0.0, -0.4))
    rng = np.random.default_rng(21)
    s = int(np.ceil(n ** (1.0 / 3.0)))
    g = np.stack(np.meshgrid(*[np.arange(s)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
    pos = np.ones((n, 4), np.float32)
    pos[:, :3] = (g * (2.4 * r) + np.array([-0.3, -0.3, -0.25]) +
                  rng.uniform(-0.1 * r, 0.1 * r, size=(n, 3)))
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.uniform(-0.002, 0.002, size=(n, 3))
    return (v, f), pos, vel


@pytest.mark.parametrize("steps", [40])
def test_system_distmap_boundary_bitexact(gpu_ctx, steps):
    from fullc0ntact_amd import ParticleSystem
    from fullc0ntact_amd.distance_map import DistanceMap
    n, grid, r = 8000, 64, 1.0 / 64
    (v, f), pos, vel = _knot_scene(n, r)
    gm = DistanceMap.from_mesh(gpu_ctx, v, f, (48, 48, 48), 2.0 / 48, (-1.0, -1.0, -1.0))
    m, _ = _oracle_from(gm)
    ps = ParticleSystem(n, (grid, grid, grid), context=gpu_ctx)
    ps.setGravity(-0.003)
    ps.setParticles(pos, vel)
    ps.setBoundary(gm)
    P = orc.params_from(ps.m_params)
    st = orc.State(ps.m_hPos, ps.m_hVel, P.numCells)
    for k in range(steps):
        ps.update(0.5)
        orc.update_distmap(st, 0.5, P, m)
    gp = ps.getArray(ps.POSITION)
    gv = ps.getArray(ps.VELOCITY)
    assert np.array_equal(bits(gp), bits(st.pos4))
    assert np.array_equal(bits(gv), bits(st.vel4))
    # the graph-launched run() of the same steps from the same start
    ps.setParticles(pos, vel)
    ps.run(0.5, steps)
    assert np.array_equal(bits(ps.getArray(ps.POSITION)), bits(st.pos4))
    assert np.array_equal(bits(ps.getArray(ps.VELOCITY)), bits(st.vel4))
    # some particles reached the map surface
    d, _, _ = orc.distmap_query(m, st.pos4[:, :3])
    assert (d < 2 * r).sum() > 0
    ps.setBoundary(None)
    ps.close()
    gm.close()


def test_distmap_argument_errors(gpu_ctx):
    from fullc0ntact_amd import _lib
    from fullc0ntact_amd.distance_map import DistanceMap
    v, f = dc.small_sphere(0.5, 8, 4)
    bad = f.copy()
    bad[3, 1] = len(v)
    with pytest.raises(_lib.FcError):
        DistanceMap.from_mesh(gpu_ctx, v, bad, (4, 4, 4), 0.5, (-1, -1, -1))
    with pytest.raises(_lib.FcError):
        DistanceMap.from_mesh(gpu_ctx, v, f, (0, 4, 4), 0.5, (-1, -1, -1))
    with pytest.raises(_lib.FcError):
        DistanceMap.from_mesh(gpu_ctx, v, f, (4, 4, 4), 0.0, (-1, -1, -1))
    m, lay, _ = dc.plane_map((4, 4, 4), 0.5, (-1, -1, -1), 0.0)
    coords = lay["vertex_coords"].copy()
    coords[7, 0] += 0.01  # not a tensor-product lattice
    with pytest.raises(_lib.FcError):
        DistanceMap.from_arrays(gpu_ctx, (4, 4, 4), 0.5, (-1, -1, -1), (1, 1, 1), coords,
                                lay["normals"], lay["contact_points"], lay["distance"])
    gm = DistanceMap.from_arrays(gpu_ctx, (4, 4, 4), 0.5, (-1, -1, -1), (1, 1, 1),
                                 lay["vertex_coords"], lay["normals"], lay["contact_points"],
                                 lay["distance"])
    d, _, _ = gm.query(np.zeros((0, 3), np.float32))
    assert d.size == 0
    back = gm.download()
    for k in lay:
        assert np.array_equal(back[k], lay[k]), k
    gm.close()

"""Pins the oracle's distance-map functions (fc_oracle_distmap.c) on the CPU.

 - point-triangle distance: against the textbook plane/edge decomposition
   (all seven regions), and against the reference's own
   CDistancePointTriangle built from its sources (oracle/_ref), which agrees
   in regions 0, 1, 2, 3, 5 and departs in 4 and 6;
 - mesh signed distance: the analytic sphere distance within the
   tessellation error, brute force over triangles, and the sign against a
   ray-parity inside test;
 - map queries: vertex values returned exactly, linear fields reproduced by
   trilinear interpolation, the 5.0 sentinel outside the box, and the
   reference's own known-answer check (test_dist_comp, cuda_addon/
   unit_tests.cuh:290-306: the query at (0.001, 0, 0) within 1e-3 relative
   of the exact value);
 - the distance-map boundary of integrateSystem (boundaries.cuh:38-89): push
   out of the map surface, the cylinder and the z planes;
 - eval_distmap with identity and rigid transforms.
"""
from __future__ import annotations

import numpy as np
import pytest

import distmap_cases as dc
from fullc0ntact_amd import meshes
from oracle import oracle as orc

RNG = np.random.default_rng(20240611)


def _triangles():
    yield np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0])
    yield np.array([0.1, -0.2, 0.3]), np.array([1.3, 0.4, -0.2]), np.array([-0.5, 0.9, 0.6])
    # obtuse and elongated
    yield np.array([0.0, 0, 0]), np.array([3.0, 0, 0]), np.array([2.5, 0.3, 0.1])
    for _ in range(3):
        yield tuple(RNG.normal(size=3) for _ in range(3))


def _points(n=3000, scale=2.5):
    return RNG.uniform(-scale, scale, size=(n, 3))


def test_point_triangle_matches_closed_form_all_regions():
    seen = set()
    for a, b, c in _triangles():
        q = _points()
        reg = dc.eberly_region(q, a, b, c)
        seen |= set(reg.tolist())
        dref, cref = dc.closed_form_point_triangle(q, a, b, c)
        for i in range(len(q)):
            d2, cp = orc.point_triangle_dist2(q[i], a, b, c)
            assert abs(np.sqrt(d2) - dref[i]) <= 1e-12 * max(1.0, dref[i]), (reg[i], q[i])
            assert np.allclose(cp, cref[i], atol=1e-10), (reg[i], q[i], cp, cref[i])
    assert seen == set(range(7))


def test_point_triangle_on_triangle_and_vertices():
    a, b, c = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0])
    for p in (a, b, c, (a + b) / 2, (b + c) / 2, (a + b + c) / 3):
        d2, cp = orc.point_triangle_dist2(p, a, b, c)
        assert d2 == 0.0 and np.array_equal(cp, p)
    d2, cp = orc.point_triangle_dist2([0.2, 0.2, 1.0], a, b, c)
    assert d2 == 1.0 and np.allclose(cp, [0.2, 0.2, 0.0])
    d2, cp = orc.point_triangle_dist2([-1.0, -1.0, 0.0], a, b, c)
    assert d2 == 2.0 and np.array_equal(cp, [0.0, 0.0, 0.0])


ref_pt = orc.reference_point_triangle()


@pytest.mark.skipif(ref_pt is None, reason="oracle/_ref not built (needs the reference tree)")
def test_point_triangle_against_reference_build():
    """CDistancePointTriangle<double>::ComputeDistance
    (distance/distancetriangle.cpp:55-190) compiled from the reference."""
    agree_regions = {0, 1, 2, 3, 5}
    differ = {4: 0, 6: 0}
    total = {4: 0, 6: 0}
    rng = np.random.default_rng(1)
    tris = [tuple(rng.normal(size=3) for _ in range(3)) for _ in range(5)]
    # a triangle where the region-6 branch condition selects the wrong edge
    tris.append((np.array([-0.85051416, -1.01887767, -0.16804871]),
                 np.array([0.24047611, -0.5168178, 0.10322873]),
                 np.array([0.90639566, 0.96987904, -0.4967267])))
    for a, b, c in tris:
        q = rng.uniform(-4.0, 4.0, size=(1500, 3))
        reg = dc.eberly_region(q, a, b, c)
        for i in range(len(q)):
            d2, cp = orc.point_triangle_dist2(q[i], a, b, c)
            dr, cr = ref_pt(q[i], a, b, c)
            if reg[i] in agree_regions:
                assert abs(np.sqrt(d2) - dr) <= 1e-12 * max(1.0, dr), (reg[i], q[i])
                assert np.allclose(cp, cr, atol=1e-10)
            else:
                total[reg[i]] += 1
                differ[reg[i]] += int(abs(np.sqrt(d2) - dr) > 1e-9)
    # the reference's region-4/6 formulas select a different point there
    assert total[4] > 0 and total[6] > 0
    assert differ[4] > 0 and differ[6] > 0


def test_mesh_sdf_sphere_analytic():
    R = 0.5
    v, f = meshes.uv_sphere(R, 48, 24)
    q = RNG.uniform(-1.0, 1.0, size=(4000, 3))
    sdf, cp, w = orc.mesh_sdf(v, f, q)
    r = np.linalg.norm(q, axis=1)
    sag = R * (1.0 - np.cos(2 * np.pi / 48))  # chord deviation bound
    assert np.all(np.abs(sdf - (r - R)) <= sag + 1e-12)
    clear = np.abs(r - R) > sag
    assert np.all(np.sign(sdf[clear]) == np.sign(r[clear] - R))
    assert np.allclose(w[r < R - sag], 1.0, atol=1e-9)
    assert np.allclose(w[r > R + 1e-9], 0.0, atol=1e-9)
    # closest points lie on the tessellated surface
    assert np.all(np.abs(np.linalg.norm(cp, axis=1) - R) <= sag + 1e-12)
    assert np.allclose(np.linalg.norm(q - cp, axis=1), np.abs(sdf), rtol=0, atol=1e-12)


def test_mesh_sdf_brute_force_and_ray_parity():
    v, f = dc.golden_mesh("hopper")
    lo, hi = v.min(0) - 0.2, v.max(0) + 0.2
    q = RNG.uniform(lo, hi, size=(300, 3))
    sdf, cp, w = orc.mesh_sdf(v, f, q)
    best = np.full(len(q), np.inf)
    # the mesh holds zero-area triangles (repeated vertices); they add no
    # distance candidates of their own, their edges belong to neighbours
    area = np.linalg.norm(np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]]), axis=1)
    assert (area == 0).sum() == 36
    for t in f[area > 0]:
        d, _ = dc.closed_form_point_triangle(q, v[t[0]], v[t[1]], v[t[2]])
        best = np.minimum(best, d)
    assert np.allclose(np.abs(sdf), best, rtol=0, atol=1e-12)
    inside = dc.ray_parity_inside(v, f, q)
    clear = np.abs(sdf) > 1e-6
    assert np.array_equal((sdf < 0)[clear], inside[clear])
    assert np.array_equal((w > 0.5), inside)


def test_mesh_sdf_torusknot_sign():
    v, f = dc.golden_mesh("torusknot1")
    lo, hi = v.min(0), v.max(0)
    q = RNG.uniform(lo, hi, size=(150, 3))
    # points on the tube centre line are inside
    sdf, _, w = orc.mesh_sdf(v, f, q)
    inside = dc.ray_parity_inside(v, f, q)
    assert np.array_equal(sdf < 0, inside)
    assert np.all((np.abs(w - 1.0) < 1e-6) | (np.abs(w) < 1e-6))


def _sphere_map(cells=(16, 16, 16), R=0.5):
    v, f = dc.small_sphere(R, 24, 12)
    cs = 1.5 / cells[0]
    return (v, f) + dc.oracle_map(v, f, cells, cs, (-0.75, -0.75, -0.75)) + (cs,)


def test_query_returns_vertex_values():
    v, f, m, lay, bmax, cs = _sphere_map()
    q = lay["vertex_coords"]
    d, cp, nr = orc.distmap_query(m, q)
    assert np.array_equal(d, lay["distance"])
    assert np.array_equal(cp, lay["contact_points"])
    # the normal is the stored normal of the base vertex of the cell
    base = [np.array_equal(nr[i], lay["normals"][i]) for i in range(len(q))]
    assert np.mean(base) > 0.8


def test_query_reproduces_linear_field():
    cells, cs, bmin = (8, 8, 8), 0.25, (-1.0, -1.0, -1.0)
    m, lay, bmax = dc.plane_map(cells, cs, bmin, z0=0.125)
    q = RNG.uniform(-1.0, 1.0, size=(5000, 3)).astype(np.float32)
    d, cp, _ = orc.distmap_query(m, q)
    assert np.allclose(d, q[:, 2] - np.float32(0.125), atol=2e-7)
    assert np.allclose(cp[:, :2], q[:, :2], atol=2e-7)
    assert np.allclose(cp[:, 2], 0.125, atol=2e-7)


def test_query_outside_sentinel_and_faces():
    cells, cs, bmin = (4, 4, 4), 0.5, (-1.0, -1.0, -1.0)
    m, lay, bmax = dc.plane_map(cells, cs, bmin, z0=0.0)
    eps = np.float32(1e-6)
    pts = np.array([[1.0 + eps, 0, 0], [-1.0 - eps, 0, 0], [0, 1.0 + eps, 0], [0, -1.0 - eps, 0],
                    [0, 0, 1.0 + eps], [0, 0, -1.0 - eps], [np.nan, 0, 0]], np.float32)
    d, _, _ = orc.distmap_query(m, pts)
    assert np.all(d == 5.0)
    # the closed box includes its faces; the upper faces use the last cell
    faces = np.array([[1.0, 0.3, 0.2], [0.3, 1.0, 0.2], [0.3, 0.2, 1.0], [-1.0, -1.0, -1.0],
                      [1.0, 1.0, 1.0]], np.float32)
    d, cp, _ = orc.distmap_query(m, faces)
    assert np.allclose(d, faces[:, 2], atol=1e-7)


def test_dist_comp_known_answer():
    """test_dist_comp (cuda_addon/unit_tests.cuh:290-306): the map's value at
    (0.001, 0, 0) within 1e-3 relative of the exact signed distance."""
    v, f, m, lay, bmax, cs = _sphere_map((32, 32, 32))
    d, _, _ = orc.distmap_query(m, np.array([[0.001, 0.0, 0.0]], np.float32))
    exact, _, _ = orc.mesh_sdf(v, f, np.array([[0.001, 0.0, 0.0]]))
    assert abs((d[0] - exact[0]) / exact[0]) < 1e-3


def _params(r=1.0 / 64):
    return orc.make_params(64, 64, r, gravity=(0.0, 0.0, 0.0), globalDamping=1.0)


def test_integrate_distmap_push_out_of_plane():
    cells, cs, bmin = (8, 8, 8), 0.25, (-1.0, -1.0, -1.0)
    m, _, _ = dc.plane_map(cells, cs, bmin, z0=0.0)
    P = _params()
    r, bd = P.particleRadius, P.boundaryDamping
    pos = np.array([[0.1, 0.2, 0.5 * r, 1.0], [0.1, 0.2, 0.5, 1.0]], np.float32)
    vel = np.array([[0.0, 0.0, 0.0, 7.0], [0.0, 0.0, 0.0, 3.0]], np.float32)
    vel[:, :3] = [0.01, -0.02, -0.001]
    p2, v2 = orc.integrate_distmap(pos, vel, 1.0, P, m)
    # first particle: below r of the plane -> pushed to z = r + 0.1 r
    assert abs(p2[0, 2] - 1.1 * r) < 1e-6
    assert np.allclose(p2[0, :2], pos[0, :2] + vel[0, :2], atol=1e-7)
    assert np.allclose(v2[0, :3], vel[0, :3] * np.array([0.5, 0.5, 0.7]) * bd, atol=1e-9)
    # second: free flight; w of both preserved
    assert np.allclose(p2[1, :3], pos[1, :3] + vel[1, :3], atol=1e-7)
    assert np.array_equal(v2[1, :3], vel[1, :3])
    assert np.array_equal(v2[:, 3], vel[:, 3]) and np.array_equal(p2[:, 3], pos[:, 3])


def test_integrate_distmap_cylinder_and_planes():
    cells, cs, bmin = (4, 4, 4), 0.25, (-0.5, -0.5, -0.5)
    m, _, _ = dc.plane_map(cells, cs, bmin, z0=-10.0)  # map far below: d >> r
    P = _params()
    r, bd = P.particleRadius, P.boundaryDamping
    pos = np.array([[1.05, 0.0, 2.0, 1], [0.0, 0.0, 10.0, 1], [0.0, 0.0, -1.0, 1]], np.float32)
    vel = np.array([[0.01, 0.0, 0.02, 0], [0.0, 0.0, 0.01, 0], [0.0, 0.0, -0.01, 0]], np.float32)
    p2, v2 = orc.integrate_distmap(pos, vel, 1.0, P, m)
    assert abs(np.hypot(p2[0, 0], p2[0, 1]) - (1.0 - r)) < 1e-6
    assert np.allclose(v2[0, :2], vel[0, :2] * bd) and v2[0, 2] == vel[0, 2]
    assert p2[1, 2] == np.float32(10.0 - r) and v2[1, 2] == np.float32(0.01) * np.float32(bd)
    assert p2[2, 2] == np.float32(-1.0 + r) and v2[2, 2] == np.float32(-0.01) * np.float32(bd)


def _trans(R0, o0, R1, o1):
    t = orc.OrcTransInfo()
    R0 = np.asarray(R0, np.float32)
    R1 = np.asarray(R1, np.float32)
    for a in range(3):
        t.origin0[a] = float(o0[a])
        t.origin1[a] = float(o1[a])
    for k in range(9):
        t.m2w0[k] = float(R0.ravel()[k])
        t.w2m0[k] = float(R0.T.ravel()[k])
        t.m2w1[k] = float(R1.ravel()[k])
    return t


def _rot(ax, ang):
    c, s = np.cos(ang), np.sin(ang)
    if ax == "z":
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def test_eval_identity_and_rigid_motion():
    v, f, m, lay, bmax, cs = _sphere_map()
    q = RNG.uniform(-0.9, 0.9, size=(2000, 3)).astype(np.float32)
    I = np.eye(3)
    cps, nr, d = orc.distmap_eval(m, q, _trans(I, (0, 0, 0), I, (0, 0, 0)))
    dq, cq, _ = orc.distmap_query(m, q)
    inb = np.all(np.abs(q) <= 0.75, axis=1)
    assert np.array_equal(d[inb], dq[inb]) and np.all(d[~inb] == 1000.0)
    assert np.allclose(cps[inb], 0.5 * (cq[inb] + q[inb]), atol=1e-6)
    assert np.all(cps[~inb] == 0.0)
    # moving both frames by the same rigid motion leaves distances unchanged
    R = _rot("z", 0.3) @ _rot("x", -0.2)
    o = np.array([0.3, -0.1, 0.2])
    qb = (q.astype(np.float64) @ R.T + o).astype(np.float32)  # body-1 model coords
    t = _trans(R, o, I, (0, 0, 0))
    _, _, d2 = orc.distmap_eval(m, qb, t)
    sel = inb & np.all(np.abs(q) <= 0.7, axis=1)
    assert np.allclose(d2[sel], dq[sel], atol=2e-3)


def test_update_distmap_matches_box_update_away_from_walls():
    """With every particle clear of the map surface, the cylinder and the
    z planes, and clear of the box walls, both boundaries leave the step a
    plain semi-implicit Euler step, so the two updates agree bit for bit."""
    n, grid, r = 512, 32, 1.0 / 64
    P = orc.make_params(n, grid, r)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = RNG.uniform(-0.4, 0.4, size=(n, 3))
    pos[:, 3] = 1.0
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = RNG.uniform(-1e-3, 1e-3, size=(n, 3))
    m, _, _ = dc.plane_map((4, 4, 4), 0.5, (-1, -1, -1), z0=-50.0)
    a = orc.State(pos, vel, P.numCells)
    b = orc.State(pos, vel, P.numCells)
    for _ in range(5):
        a.update(0.5, P)
        orc.update_distmap(b, 0.5, P, m)
    assert np.array_equal(a.pos4, b.pos4) and np.array_equal(a.vel4, b.vel4)

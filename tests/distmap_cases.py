"""Shared cases for the distance-map tests: meshes, lattices, maps built by
the oracle, and independent geometric checkers (closed-form point-triangle
distance, ray-parity inside test) that do not share code with the oracle."""
from __future__ import annotations

from pathlib import Path

import numpy as np

from fullc0ntact_amd import meshes
from oracle import oracle as orc

GOLDEN = Path(__file__).resolve().parent / "golden"


def golden_mesh(name: str):
    """Vertices/triangles of a reference mesh file (hopper, torusknot1)."""
    with np.load(GOLDEN / "meshes.npz") as z:
        return z[f"{name}_vertices"], z[f"{name}_triangles"]


def lattice_points(cells, cell_size, bmin):
    """Lattice vertices bmin + i*cellSize in binary64, x fastest."""
    cx, cy, cz = cells
    iz, iy, ix = np.meshgrid(np.arange(cz + 1), np.arange(cy + 1), np.arange(cx + 1),
                             indexing="ij")
    b = np.asarray(bmin, np.float32).astype(np.float64)
    cs = float(np.float32(cell_size))
    return np.stack([b[0] + ix.ravel() * cs, b[1] + iy.ravel() * cs, b[2] + iz.ravel() * cs], 1)


def oracle_map(verts, tris, cells, cell_size, bmin):
    """The map fc_distmap_build produces, computed by the oracle:
    (DistMap, reference-layout dict)."""
    q = lattice_points(cells, cell_size, bmin)
    sdf, cp, _ = orc.mesh_sdf(verts, tris, q)
    bmin32 = np.asarray(bmin, np.float32)
    cs = np.float32(cell_size)
    bmax32 = np.array([np.float32(float(bmin32[a]) + cells[a] * float(cs)) for a in range(3)],
                      np.float32)
    layout = {
        "vertex_coords": q.astype(np.float32),
        "normals": (q - cp).astype(np.float32),
        "contact_points": cp.astype(np.float32),
        "distance": sdf.astype(np.float32),
    }
    m = orc.DistMap.from_reference_layout(cells, cs, bmin32, bmax32, layout["vertex_coords"],
                                          layout["normals"], layout["contact_points"],
                                          layout["distance"])
    return m, layout, bmax32


def plane_map(cells, cell_size, bmin, z0):
    """Reference-layout map of the half space z > z0: distance z - z0, closest
    point (x, y, z0), normal (0, 0, z - z0)."""
    q = lattice_points(cells, cell_size, bmin).astype(np.float32)
    cp = q.copy()
    cp[:, 2] = np.float32(z0)
    d = (q[:, 2] - np.float32(z0)).astype(np.float32)
    nrm = np.zeros_like(q)
    nrm[:, 2] = d
    bmin32 = np.asarray(bmin, np.float32)
    bmax32 = np.array([np.float32(float(bmin32[a]) + cells[a] * float(np.float32(cell_size)))
                       for a in range(3)], np.float32)
    layout = {"vertex_coords": q, "normals": nrm, "contact_points": cp, "distance": d}
    m = orc.DistMap.from_reference_layout(cells, np.float32(cell_size), bmin32, bmax32, q, nrm,
                                          cp, d)
    return m, layout, bmax32


def closed_form_point_triangle(q, a, b, c):
    """Distance from points q (n,3) to triangle abc by the textbook
    decomposition: if the projection onto the plane falls inside the
    triangle, the plane distance; else the least of the three point-segment
    distances.  Returns (dist, closest point)."""
    q = np.atleast_2d(q).astype(np.float64)
    a, b, c = (np.asarray(x, np.float64) for x in (a, b, c))
    n = np.cross(b - a, c - a)
    nn = n / np.linalg.norm(n)
    h = (q - a) @ nn
    p = q - h[:, None] * nn
    # barycentric inside test via edge-normal signs
    inside = np.ones(len(q), bool)
    for u, v in ((a, b), (b, c), (c, a)):
        inside &= (np.cross(v - u, p - u) @ n) >= 0.0

    def seg(u, v):
        t = np.clip(((q - u) @ (v - u)) / ((v - u) @ (v - u)), 0.0, 1.0)
        x = u + t[:, None] * (v - u)
        return np.linalg.norm(q - x, axis=1), x

    cands = [seg(a, b), seg(b, c), seg(c, a)]
    ds = np.stack([cd[0] for cd in cands], 1)
    k = np.argmin(ds, 1)
    best = ds[np.arange(len(q)), k]
    bx = np.stack([cd[1] for cd in cands], 1)[np.arange(len(q)), k]
    dist = np.where(inside, np.abs(h), best)
    cp = np.where(inside[:, None], p, bx)
    return dist, cp


def eberly_region(q, a, b, c):
    """Region (0..6) of Eberly's parameter-plane partition for each point."""
    q = np.atleast_2d(q).astype(np.float64)
    e0, e1 = b - a, c - a
    D = a - q
    A, B, Cc = e0 @ e0, e0 @ e1, e1 @ e1
    d, e = D @ e0, D @ e1
    det = A * Cc - B * B
    s, t = B * e - Cc * d, B * d - A * e
    reg = np.empty(len(q), int)
    lo = s + t <= det
    reg[lo & (s < 0) & (t < 0)] = 4
    reg[lo & (s < 0) & (t >= 0)] = 3
    reg[lo & (s >= 0) & (t < 0)] = 5
    reg[lo & (s >= 0) & (t >= 0)] = 0
    reg[~lo & (s < 0)] = 2
    reg[~lo & (s >= 0) & (t < 0)] = 6
    reg[~lo & (s >= 0) & (t >= 0)] = 1
    return reg


def ray_parity_inside(verts, tris, pts, direction=(0.5773, 0.5774, 0.5775)):
    """Inside test by counting crossings of a ray (Moller-Trumbore) with the
    mesh: odd = inside.  An independent check of a winding-number sign."""
    dvec = np.asarray(direction, np.float64)
    dvec = dvec / np.linalg.norm(dvec)
    v0, v1, v2 = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    e1, e2 = v1 - v0, v2 - v0
    h = np.cross(dvec, e2)
    det = np.einsum("ij,ij->i", e1, h)
    ok = np.abs(det) > 1e-14
    inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
    out = np.zeros(len(pts), bool)
    for i, p in enumerate(np.asarray(pts, np.float64)):
        s = p - v0
        u = np.einsum("ij,ij->i", s, h) * inv
        qv = np.cross(s, e1)
        v = (qv @ dvec) * inv
        t = np.einsum("ij,ij->i", e2, qv) * inv
        hit = ok & (u >= 0) & (v >= 0) & (u + v <= 1) & (t > 0)
        out[i] = hit.sum() % 2 == 1
    return out


def small_sphere(radius=0.5, nu=24, nv=12, center=(0.0, 0.0, 0.0)):
    return meshes.uv_sphere(radius, nu, nv, center)

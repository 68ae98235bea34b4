"""Distance-map files on the GPU: load, save, FBM states and queryFBM.

 - a map built from a mesh, saved and loaded again with its box, is the
   same map bit for bit (nine significant digits round-trip binary32), with
   bit-identical queries;
 - a reference-style file (six significant digits) loads as the binary32
   rounding of the parsed binary64 values (the reference's conversion
   constructor), and its queries equal the oracle's on those arrays; without
   a box the first and last lattice vertices span it;
 - FBM states travel with the map; queryFBM is 1 exactly where the oracle's
   interpolated distance inside the box is <= 0.
"""
from __future__ import annotations

import numpy as np
import pytest

import distmap_cases as dc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _queries(bmin, bmax, n=20000, seed=3):
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(bmin), np.asarray(bmax)
    pad = 0.1 * (hi - lo)
    return rng.uniform(lo - pad, hi + pad, (n, 3)).astype(np.float32)


def test_save_load_round_trip_is_exact(gpu_ctx, tmp_path):
    from fullc0ntact_amd.distance_map import DistanceMap
    v, t = dc.small_sphere(0.5, 24, 12)
    m = DistanceMap.from_mesh(gpu_ctx, v, t, (20, 24, 16), 1.5 / 20, (-0.75, -0.8, -0.7))
    states = (np.arange(m.num_vertices) % 3 == 0).astype(np.int32)
    m.set_fbm(states)
    path = tmp_path / "sphere.dmap"
    m.save(path)
    m2 = DistanceMap.load(gpu_ctx, path, m.bmin, m.bmax)
    a, b = m.download(), m2.download()
    assert m2.cells == m.cells and m2.cell_size == m.cell_size
    for k in a:
        assert np.array_equal(bits(a[k]), bits(b[k])), k
    assert np.array_equal(m2.fbm(), states)
    q = _queries(m.bmin, m.bmax)
    for x, y in zip(m.query(q), m2.query(q)):
        assert np.array_equal(bits(x), bits(y))
    m.close()
    m2.close()


def test_reference_style_file_matches_oracle(gpu_ctx, tmp_path):
    from fullc0ntact_amd.distance_map import DistanceMap, read_map_file, write_map_file
    verts, tris = dc.golden_mesh("hopper")
    cells, cs, bmin = (14, 14, 10), 0.15, (-1.05, -1.05, -0.55)
    om, lay, _ = dc.oracle_map(verts, tris, cells, cs, bmin)
    fbm = (lay["distance"] <= 0).astype(np.int32)
    path = tmp_path / "hopper.dmap"
    write_map_file(path, cells, cs, lay["vertex_coords"], lay["contact_points"],
                   lay["normals"], lay["distance"], fbm, digits=6)
    host = read_map_file(path)
    m = DistanceMap.load(gpu_ctx, path)  # box from the first and last vertex
    f32 = {k: host[k].astype(np.float32) for k in ("vertex_coords", "contact_points",
                                                     "normals", "distance")}
    assert m.cell_size == np.float32(host["cell_size"])
    assert np.array_equal(np.float32(m.bmin), f32["vertex_coords"][0])
    assert np.array_equal(np.float32(m.bmax), f32["vertex_coords"][-1])
    got = m.download()
    for k in f32:
        assert np.array_equal(bits(got[k]), bits(f32[k].reshape(got[k].shape))), k
    assert np.array_equal(m.fbm(), fbm)
    ref = orc.DistMap.from_reference_layout(cells, np.float32(host["cell_size"]),
                                            f32["vertex_coords"][0], f32["vertex_coords"][-1],
                                            f32["vertex_coords"], f32["normals"],
                                            f32["contact_points"], f32["distance"])
    q = _queries(m.bmin, m.bmax, seed=9)
    gd, gc, gn = m.query(q)
    rd, rc, rn = orc.distmap_query(ref, q)
    assert np.array_equal(bits(gd), bits(rd))
    assert np.array_equal(bits(gc), bits(rc)) and np.array_equal(bits(gn), bits(rn))
    inside = m.query_fbm(q)
    assert np.array_equal(inside, (rd <= 0).astype(np.int32))
    assert 0 < inside.sum() < len(q)
    m.close()

"""The learned signed-distance workload at its stated size: 2,097,152 spheres
of radius 1/256 around 64 rigid bodies on a 4^3 lattice of [-1, 1]^3, every
body its own seeded shape (torus knot or superellipsoid) with its own
256^3 distance map and its own network (32 Fourier features -> 3 x 64 SiLU
-> 1, BF16 MFMA inference with FP32 accumulation) trained on 2^20 exact
mesh distances.

 - the exact path's contact flags equal the oracle's queryMap
   (distancemap.cuh:154-191) under checkDistMap's contact test
   (boundaries.cuh:38-63) on every sphere assigned to a sample of bodies;
 - the surrogate-culled contact set equals the exact one on all 2M spheres,
   with each network's margin derived from its map's lattice
   (boundary.lattice_margins), and most exact queries are culled.
"""
from __future__ import annotations

import os
import time

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

N_SPHERES = 1 << 21
RADIUS = 1.0 / 256
SAMPLES = 1 << 20


@pytest.fixture(scope="module", autouse=True)
def _threads():
    orc.set_threads(max(1, min(16, os.cpu_count() or 1)))
    yield
    orc.set_threads(1)


@pytest.fixture(scope="module")
def workload(gpu_ctx):
    from fullc0ntact_amd import _lib
    from fullc0ntact_amd.boundary import build_scene
    t0 = time.perf_counter()
    print("[sdf-workload] 64 shapes, 256^3 maps, 2^20 samples per network", flush=True)
    sc = build_scene(gpu_ctx, lattice=4, cells=256, samples=SAMPLES, epochs=30, batch=16384,
                     seed=1973, radius=RADIUS, shapes=None, knot_segments=(128, 8),
                     ellipsoid_segments=(32, 16),
                     log=lambda m: print(f"[sdf-workload] {m}", flush=True))
    print(f"[sdf-workload] maps {sc.build_s:.1f} s, training + margins {sc.train_s:.1f} s "
          f"(total {time.perf_counter() - t0:.1f} s)", flush=True)
    # spheres in the free space around the bodies: uniform draws whose
    # centre lies inside a body (map distance < 0) are dropped
    rng = np.random.default_rng(2024)
    draw = np.ones((N_SPHERES + N_SPHERES // 4, 4), np.float32)
    draw[:, :3] = rng.uniform(-1.0, 1.0, (len(draw), 3))
    dd = _lib.DeviceArray.from_host(gpu_ctx, draw)
    inside, _ = sc.bodies.contacts(dd, len(draw), 0.0)
    keep = inside.download(np.uint8, len(draw)) == 0
    pos = np.ascontiguousarray(draw[keep][:N_SPHERES])
    assert len(pos) == N_SPHERES
    dd.free()
    dp = _lib.DeviceArray.from_host(gpu_ctx, pos)
    yield sc, pos, dp
    dp.free()
    sc.bodies.close()
    sc.surrogate.close()
    for m in sc.shape_maps:
        m.close()


def _oracle_flags_of_bodies(sc, pos, radius, bodies):
    """Oracle contact flags of the spheres whose lattice cell holds one of
    `bodies`, with the kernel's float32 arithmetic for the cell and the
    model-frame point; returns (sphere indices, flags, in-box mask)."""
    bs = sc.bodies
    p = pos[:, :3].astype(np.float32)
    cq = (p - bs.lattice_origin) / bs.lattice_cell
    cidx = np.floor(cq).astype(np.int64)  # every centre is finite here
    lat = bs.lattice
    ok = np.all((cidx >= 0) & (cidx < lat), axis=1)
    cell = (cidx[:, 2] * lat[1] + cidx[:, 1]) * lat[0] + cidx[:, 0]
    body = np.where(ok, bs.cell_body[np.clip(cell, 0, len(bs.cell_body) - 1)], -1)
    idx, flags, boxes = [], [], []
    for b in bodies:
        sel = np.nonzero(body == b)[0]
        w2m = bs.m2w[b].reshape(3, 3).T.astype(np.float32)
        d = (p[sel] - bs.origins[b]).astype(np.float32)
        q = np.zeros_like(d)
        for i in range(3):
            s = np.zeros(len(d), np.float32)
            s = s + w2m[i, 0] * d[:, 0]
            s = s + w2m[i, 1] * d[:, 1]
            s = s + w2m[i, 2] * d[:, 2]
            q[:, i] = s
        gm = sc.maps[b]
        lay = gm.download()
        m = orc.DistMap.from_reference_layout(gm.cells, np.float32(gm.cell_size),
                                              np.asarray(gm.bmin, np.float32),
                                              np.asarray(gm.bmax, np.float32),
                                              lay["vertex_coords"], lay["normals"],
                                              lay["contact_points"], lay["distance"])
        del lay
        dist, _, _ = orc.distmap_query(m, q)
        del m
        box = np.all((q >= np.asarray(gm.bmin, np.float32)) &
                     (q <= np.asarray(gm.bmax, np.float32)), axis=1)
        idx.append(sel)
        flags.append((box & (dist < np.float32(radius))).astype(np.uint8))
        boxes.append(box)
    return np.concatenate(idx), np.concatenate(flags), np.concatenate(boxes)


def test_exact_path_matches_oracle_query_map(gpu_ctx, workload):
    sc, pos, dp = workload
    assert sc.bodies.num_bodies == 64 and len(sc.shape_maps) == 64
    assert all(m.cells == (256, 256, 256) for m in sc.shape_maps)
    exact, cnt = sc.bodies.contacts(dp, N_SPHERES, RADIUS)
    fe = exact.download(np.uint8, N_SPHERES)
    print(f"[sdf-workload] in box {cnt[0]}, contacts {cnt[2]}", flush=True)
    assert cnt[2] == fe.sum() > 10_000
    sample = [0, 9, 22, 37, 50, 63]  # knots and superellipsoids, corners and interior
    idx, ref, box = _oracle_flags_of_bodies(sc, pos, RADIUS, sample)
    print(f"[sdf-workload] oracle: {len(idx)} spheres of bodies {sample}, {box.sum()} in box, "
          f"{ref.sum()} contacts", flush=True)
    assert np.array_equal(fe[idx], ref)
    assert ref.sum() > 500


def test_surrogate_contact_set_equals_exact(gpu_ctx, workload):
    sc, pos, dp = workload
    exact, cnt_e = sc.bodies.contacts(dp, N_SPHERES, RADIUS)
    fe = exact.download(np.uint8, N_SPHERES)
    culled, cnt_s = sc.bodies.contacts(dp, N_SPHERES, RADIUS, surrogate=sc.surrogate,
                                       margins=sc.margins)
    fs = culled.download(np.uint8, N_SPHERES)
    st = sc.margin_stats
    print(f"[sdf-workload] exact queries {cnt_s[1]} of {cnt_s[0]} in box, contacts "
          f"{cnt_s[2]}; margins {np.min(sc.margins):.5f}..{np.max(sc.margins):.5f}, "
          f"largest |net - map| on the band vertices {max(s['max_err'] for s in st):.5f}, "
          f"largest edge slope {max(s['slope'] for s in st):.3f}", flush=True)
    assert np.array_equal(fs, fe)  # the same contact set, sphere for sphere
    assert cnt_s[0] == cnt_e[0] and cnt_s[2] == cnt_e[2]
    assert cnt_s[1] < 0.5 * cnt_s[0]  # most exact queries culled

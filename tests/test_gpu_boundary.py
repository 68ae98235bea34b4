"""Boundary contacts of spheres with a set of rigid bodies (distance maps in
rotated model frames), exact and surrogate-culled, against the oracle.

 - exact path: every sphere's flag equals the oracle's queryMap of the
   model-frame centre (< radius), bit for bit in the query;
 - surrogate path (BF16 MFMA network cull, exact query in the band): the
   same contact set, with most exact queries removed -- what the trained
   network gains over the method without it.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _threads():
    orc.set_threads(max(1, min(16, os.cpu_count() or 1)))
    yield
    orc.set_threads(1)


@pytest.fixture(scope="module", params=[None, 3], ids=["shape-per-body", "3-shapes"])
def scene(gpu_ctx, request):
    from fullc0ntact_amd.boundary import build_scene
    print(f"building an 8-body scene, shapes={request.param}", flush=True)
    # coarse 40^3 maps around bodies that nearly fill their boxes: the
    # lattice-derived margins would send most spheres to the exact query,
    # so this scene uses the held-out calibration (the full-size workload
    # test checks the lattice-derived ones)
    sc = build_scene(gpu_ctx, lattice=2, cells=40, samples=1 << 15, epochs=40, seed=11,
                     radius=1.0 / 64, shapes=request.param, margins="held_out")
    yield sc
    sc.bodies.close()
    sc.surrogate.close()
    for m in sc.shape_maps:
        m.close()


def _oracle_flags(sc, pos, radius):
    """Body of each sphere's lattice cell and the exact map test, with the
    kernel's float32 arithmetic for the cell and the model-frame point."""
    bs = sc.bodies
    p = pos[:, :3].astype(np.float32)
    lat = bs.lattice
    # a sphere with a non-finite coordinate is in no box (its model-frame
    # point is not finite, so the kernel's box test fails); it is masked
    # before the float -> int conversion, which numpy leaves undefined
    cq = (p - bs.lattice_origin) / bs.lattice_cell
    finite = np.all(np.isfinite(cq), axis=1)
    cidx = np.floor(np.where(finite[:, None], cq, -1.0)).astype(np.int64)
    ok = finite & np.all((cidx >= 0) & (cidx < lat), axis=1)
    flags = np.zeros(len(p), np.uint8)
    inbox = np.zeros(len(p), bool)
    cell = (cidx[:, 2] * lat[1] + cidx[:, 1]) * lat[0] + cidx[:, 0]
    body = np.where(ok, bs.cell_body[np.clip(cell, 0, len(bs.cell_body) - 1)], -1)
    for b in range(bs.num_bodies):
        sel = np.nonzero(body == b)[0]
        if sel.size == 0:
            continue
        w2m = bs.m2w[b].reshape(3, 3).T.astype(np.float32)
        d = (p[sel] - bs.origins[b]).astype(np.float32)
        q = np.zeros_like(d)
        for i in range(3):
            s = np.zeros(len(d), np.float32)
            s = s + w2m[i, 0] * d[:, 0]
            s = s + w2m[i, 1] * d[:, 1]
            s = s + w2m[i, 2] * d[:, 2]
            q[:, i] = s
        lay = sc.maps[b].download()
        m = orc.DistMap.from_reference_layout(sc.maps[b].cells, np.float32(sc.maps[b].cell_size),
                                              np.asarray(sc.maps[b].bmin, np.float32),
                                              np.asarray(sc.maps[b].bmax, np.float32),
                                              lay["vertex_coords"], lay["normals"],
                                              lay["contact_points"], lay["distance"])
        dist, _, _ = orc.distmap_query(m, q)
        box = np.all((q >= np.asarray(sc.maps[b].bmin, np.float32)) &
                     (q <= np.asarray(sc.maps[b].bmax, np.float32)), axis=1)
        inbox[sel] = box
        flags[sel] = (box & (dist < np.float32(radius))).astype(np.uint8)
    return flags, inbox


def test_exact_and_culled_contacts(gpu_ctx, scene):
    from fullc0ntact_amd import _lib
    rng = np.random.default_rng(8)
    n, radius = 300_000, 1.0 / 64
    pos = np.ones((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-1.0, 1.0, (n, 3))
    pos[:7, :3] = [[1.2, 0, 0], [-1.0, -1.0, -1.0], [1.0, 1.0, 1.0], [0, 0, 0], [0.5, 0.5, 0.5],
                   [-0.5, 0.5, -0.5], [np.nan, 0, 0]]
    dp = _lib.DeviceArray.from_host(gpu_ctx, pos)
    exact, cnt_e = scene.bodies.contacts(dp, n, radius)
    fe = exact.download(np.uint8, n)
    ref, inbox = _oracle_flags(scene, pos, radius)
    assert np.array_equal(fe, ref)
    assert cnt_e[0] == inbox.sum() and cnt_e[1] == cnt_e[0] and cnt_e[2] == ref.sum() > 100
    culled, cnt_s = scene.bodies.contacts(dp, n, radius, surrogate=scene.surrogate,
                                          margins=scene.margins)
    fs = culled.download(np.uint8, n)
    print(f"in box {cnt_s[0]}, exact queries {cnt_s[1]}, contacts {cnt_s[2]}, "
          f"margins {np.round(scene.margins, 4)}", flush=True)
    assert np.array_equal(fs, fe)  # the same contact set
    assert cnt_s[0] == cnt_e[0] and cnt_s[2] == cnt_e[2]
    assert cnt_s[1] < 0.5 * cnt_s[0]  # most exact queries culled

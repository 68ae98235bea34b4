"""Rigid-body boundaries of a particle bed: per-body distance maps and poses
(BodySet), the exact contact query and the surrogate-culled one, and the
construction of seeded multi-body scenes (torus knots and superellipsoids on
a lattice, each with its map and its trained distance network)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib, meshes
from ._lib import DeviceArray, check, lib
from .distance_map import DistanceMap, mesh_signed_distance
from .sdf import BF16, SdfSurrogate, training_samples


class BodySet:
    def __init__(self, ctx: _lib.Context, maps, model2world, origins, lattice, lattice_origin,
                 lattice_cell: float, cell_body, body_net=None):
        self.ctx = ctx
        self.maps = list(maps)
        B = len(self.maps)
        handles = (C.c_void_p * B)(*[m.handle.value for m in self.maps])
        net = np.ascontiguousarray(body_net if body_net is not None else np.arange(B), np.int32)
        m2w = np.ascontiguousarray(model2world, np.float32).reshape(B, 9)
        org = np.ascontiguousarray(origins, np.float32).reshape(B, 3)
        lat = np.ascontiguousarray(lattice, np.int32).reshape(3)
        lo = np.ascontiguousarray(lattice_origin, np.float32).reshape(3)
        cb = np.ascontiguousarray(cell_body, np.int32).reshape(-1)
        h = C.c_void_p()
        check(lib.fc_body_set_create(ctx.handle, B, handles, m2w.ctypes.data, org.ctypes.data,
                                     net.ctypes.data, lat.ctypes.data, lo.ctypes.data,
                                     C.c_float(lattice_cell), cb.ctypes.data, C.byref(h)))
        self.body_net = net
        self.handle = h
        ctx.register(self)
        self.num_bodies = B
        self.m2w, self.origins = m2w, org
        self.lattice, self.lattice_origin, self.lattice_cell, self.cell_body = lat, lo, \
            np.float32(lattice_cell), cb

    def close(self) -> None:
        if getattr(self, "handle", None):
            check(lib.fc_body_set_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def contacts(self, pos_dev: DeviceArray, n: int, radius: float, surrogate=None,
                 margins=None, candidate_dev=None, contact_dev=None, counts: bool = True):
        """Contact flags (device u8 array) and, with counts, the host
        counts (spheres in a box, exact queries, contacts)."""
        contact_dev = contact_dev or DeviceArray(self.ctx, max(n, 1))
        if surrogate is not None and candidate_dev is None:
            candidate_dev = DeviceArray(self.ctx, max(n, 1))
        m = None
        if margins is not None:
            m = np.ascontiguousarray(margins, np.float32)
        cnt = np.zeros(3, np.uint32)
        check(lib.fc_boundary_contacts(self.ctx.handle, self.handle,
                                       surrogate.handle if surrogate is not None else None,
                                       m.ctypes.data if m is not None else None, pos_dev.ptr, n,
                                       C.c_float(radius),
                                       candidate_dev.ptr if candidate_dev is not None else None,
                                       contact_dev.ptr, cnt.ctypes.data if counts else None))
        return contact_dev, (cnt if counts else None)


def rotation(rng) -> np.ndarray:
    """A uniformly random rotation matrix (QR of a Gaussian matrix)."""
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q *= np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


@dataclass
class Scene:
    bodies: BodySet
    surrogate: SdfSurrogate
    margins: np.ndarray  # per body
    meshes: list         # per shape
    maps: list           # per body (shared by the bodies of a shape)
    frames: np.ndarray   # per shape: center xyz, halfExtent (model frame)
    shape_maps: list = None
    build_s: float = 0.0
    train_s: float = 0.0
    margin_stats: list = None  # per shape, from lattice_margins


def body_mesh(kind: str, rng, size: float):
    """A seeded torus-knot tube or superellipsoid of about `size` extent,
    centred at the model origin."""
    if kind == "knot":
        p, q = [(2, 3), (3, 2), (2, 5), (3, 4)][rng.integers(0, 4)]
        v, f = meshes.torus_knot(p, q, 256, 12, radius=0.05 * size / 0.3,
                                 scale=size / 3.2 * rng.uniform(0.8, 1.0))
    else:
        a, b, c = size * rng.uniform(0.25, 0.45, 3)
        v, f = meshes.superellipsoid(a, b, c, rng.uniform(0.3, 1.2), rng.uniform(0.3, 1.2), 48, 24)
    return v, f


def build_scene(ctx: _lib.Context, lattice: int = 4, cells: int = 64, samples: int = 1 << 16,
                epochs: int = 30, seed: int = 1973, domain=(-1.0, 1.0),
                radius: float = 1.0 / 256, shapes: int | None = None,
                knot_segments=(256, 12), ellipsoid_segments=(48, 24), batch: int = 1024,
                margins: str = "lattice", log=None) -> Scene:
    """lattice^3 rigid bodies, one per lattice cell of [domain]^3, with
    random rotations, instancing `shapes` seeded shapes (torus knots and
    superellipsoids; default one shape per body): per shape a cells^3
    distance map over its box and a surrogate network trained on `samples`
    exact mesh distances (minibatches of `batch` per body); per-body margins
    from the map lattice (lattice_margins, the default) or from held-out
    points (calibrate_margins, margins="held_out")."""
    import time
    rng = np.random.default_rng(seed)
    lo, hi = domain
    cell = (hi - lo) / lattice
    B = lattice ** 3
    S = B if shapes is None else int(shapes)
    t0 = time.perf_counter()
    mesh_list, shape_maps, frames = [], [], []
    for s in range(S):
        if s % 2 == 0:
            p, q = [(2, 3), (3, 2), (2, 5), (3, 4)][rng.integers(0, 4)]
            size = 0.6 * cell
            v, f = meshes.torus_knot(p, q, knot_segments[0], knot_segments[1],
                                     radius=0.05 * size / 0.3,
                                     scale=size / 3.2 * rng.uniform(0.8, 1.0))
        else:
            a3 = 0.6 * cell * rng.uniform(0.25, 0.45, 3)
            v, f = meshes.superellipsoid(a3[0], a3[1], a3[2], rng.uniform(0.3, 1.2),
                                         rng.uniform(0.3, 1.2), *ellipsoid_segments)
        mesh_list.append((v, f))
        mn, mx = v.min(0), v.max(0)
        c = 0.5 * (mn + mx)
        half = 0.5 * float(np.max(mx - mn)) * 1.15
        cs = 2.0 * half / cells
        shape_maps.append(DistanceMap.from_mesh(ctx, v, f, (cells, cells, cells), cs, c - half))
        frames.append([*c, half])
    ctx.sync()
    build_s = time.perf_counter() - t0
    if log:
        log(f"{S} distance maps of {cells}^3 cells built in {build_s:.1f} s")
    frames = np.array(frames, np.float32)
    m2w, origins, body_net = [], [], []
    for b in range(B):
        iz, iy, ix = b // (lattice * lattice), (b // lattice) % lattice, b % lattice
        origins.append([lo + (ix + 0.5) * cell, lo + (iy + 0.5) * cell, lo + (iz + 0.5) * cell])
        m2w.append(rotation(rng).reshape(-1))
        body_net.append(b % S)
    maps = [shape_maps[k] for k in body_net]
    bodies = BodySet(ctx, maps, np.array(m2w), np.array(origins), (lattice,) * 3, (lo,) * 3,
                     cell, np.arange(B, dtype=np.int32), np.array(body_net, np.int32))
    t0 = time.perf_counter()
    sur = SdfSurrogate(ctx, frames[:, :3], frames[:, 3], freq_sigma=1.0, seed=seed)
    pts = np.zeros((S, samples, 3), np.float32)
    lab = np.zeros((S, samples), np.float32)
    for s, (v, f) in enumerate(mesh_list):
        p = training_samples(v, f, samples, frames[s, :3], float(frames[s, 3]), 0.6, 0.04,
                             seed=seed + 17 * s)
        sd, _ = mesh_signed_distance(ctx, v, f, p)
        pts[s] = p
        lab[s] = sd
    if log:
        log(f"{S} x {samples} exact-distance labels in {time.perf_counter() - t0:.1f} s; "
            f"training {S} networks, {epochs} epochs of minibatches of {batch}")
    t1 = time.perf_counter()
    dp, dl = DeviceArray.from_host(ctx, pts), DeviceArray.from_host(ctx, lab)
    del pts, lab
    sur.train(dp, dl, samples, batch, epochs, 3e-3, 0.97)
    dp.free()
    dl.free()
    ctx.sync()
    if log:
        log(f"training {time.perf_counter() - t1:.1f} s")
    t1 = time.perf_counter()
    if margins == "lattice":
        shape_margins, stats = lattice_margins(sur, shape_maps, frames, radius)
    else:
        shape_margins = calibrate_margins(ctx, sur, shape_maps, mesh_list, frames, radius, rng)
        stats = None
    ctx.sync()
    if log:
        log(f"margins ({margins}) {time.perf_counter() - t1:.1f} s")
    train_s = time.perf_counter() - t0
    body_margins = shape_margins[np.array(body_net)]
    return Scene(bodies, sur, body_margins, mesh_list, maps, frames, shape_maps, build_s, train_s,
                 stats)


def lattice_margins(sur: SdfSurrogate, maps, frames, radius: float, safety: float = 1.25):
    """Per shape k, a margin m_k with which the cull (network distance y <=
    radius + m_k) keeps every sphere whose map distance is below radius.

    The map distance d_map at a query q is the trilinear mix of its cell's
    corner values.  Along each axis its slope is a weighted mean of corner
    differences over one lattice spacing h, at most 1 because the vertex
    values are exact signed distances (1-Lipschitz); so from q to the
    nearest corner v (at most h/2 per axis) d_map changes by at most 1.5 h,
    and a contact (d_map(q) < r) has d(v) < r + 1.5 h.  Then
      y(q) - r <= (y(v) - r) + L |q - v| <= over + L h sqrt(3) / 2,
    where `over` is the largest y(v) - r over the vertices with
    d(v) < r + 1.5 h, evaluated by the cull's own BF16 inference
    (fc_sdf_map_margin), and L bounds the network's slope.  L is taken as
    `safety` times the largest secant slope along the lattice's x, y and z
    edges at that band: the margin is a bound as long as the network does
    not vary inside a lattice cell by more than `safety` times its steepest
    sampled edge.  A small absolute term covers the float rounding of the
    map query."""
    out = np.zeros(len(maps), np.float32)
    stats = []
    for k, m in enumerate(maps):
        h = float(m.cell_size)
        st = sur.map_margin(k, m, radius, 1.5 * h)
        lip = safety * st["slope"]
        out[k] = max(st["over"], 0.0) + lip * h * float(np.sqrt(3.0)) / 2.0 \
            + 1e-6 * float(frames[k, 3])
        stats.append(dict(st, h=h, margin=float(out[k])))
    return out, stats


def calibrate_margins(ctx, sur: SdfSurrogate, maps, mesh_list, frames, radius: float, rng,
                      n: int = 1 << 16, safety: float = 1.5) -> np.ndarray:
    """Per body: the margin m such that the cull (network distance <=
    radius + m) keeps every held-out point whose map distance is < radius,
    times `safety`: m = safety * max(network - radius) over those points.
    Held-out points are drawn like the training set (near the surface and
    uniform in the box) with a different stream."""
    out = np.zeros(len(maps), np.float32)
    for b, m in enumerate(maps):
        c, h = frames[b, :3], float(frames[b, 3])
        v, f = mesh_list[b]
        q = training_samples(v, f, n, c, h, 0.6, 0.04, seed=int(rng.integers(1 << 30)))
        q = q.astype(np.float32)
        d_map, _, _ = m.query(q)
        d_net = sur.eval(q, np.full(n, b, np.uint32), BF16)
        hit = d_map < radius
        over = (d_net - radius)[hit]
        out[b] = safety * max(float(over.max()) if over.size else 0.0, 0.0) + 1e-4 * h
    return out

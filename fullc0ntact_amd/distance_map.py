"""DistanceMap: a signed distance map of a rigid boundary on the GPU.

Mirrors the query side of DistanceMap<float,gpu> (cuda_addon/distancemap.cuh)
and the batch entry points of difi (cuda_addon/difi.cu): maps are uploaded
from the reference's host layout or built from a triangle mesh, then queried
by the particle boundary or in batches.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import DeviceArray, check, lib


def _i3(v):
    return (C.c_int32 * 3)(*[int(x) for x in v])


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


class DistanceMap:
    def __init__(self, ctx: _lib.Context, handle: C.c_void_p):
        self.ctx = ctx
        self.handle = handle
        ctx.register(self)
        cells = (C.c_int32 * 3)()
        cs = C.c_float()
        bmin = (C.c_float * 3)()
        bmax = (C.c_float * 3)()
        check(lib.fc_distmap_info(handle, cells, C.byref(cs), bmin, bmax))
        self.cells = tuple(cells)
        self.cell_size = cs.value
        self.bmin = tuple(bmin)
        self.bmax = tuple(bmax)
        self.num_vertices = int(np.prod([c + 1 for c in self.cells]))

    # -- construction -------------------------------------------------------
    @classmethod
    def from_arrays(cls, ctx, cells, cell_size, bmin, bmax, vertex_coords, normals,
                    contact_points, distance) -> "DistanceMap":
        """Upload a map in the reference's host layout (transferData)."""
        arrs = [np.ascontiguousarray(a, np.float32).reshape(-1)
                for a in (vertex_coords, normals, contact_points, distance)]
        h = C.c_void_p()
        check(lib.fc_distmap_create(ctx.handle, _i3(cells), C.c_float(cell_size), _f3(bmin),
                                    _f3(bmax), *[a.ctypes.data for a in arrs], C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_mesh(cls, ctx, vertices, triangles, cells, cell_size, bmin) -> "DistanceMap":
        """Exact signed distance map of a closed triangle mesh."""
        v = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
        h = C.c_void_p()
        check(lib.fc_distmap_build(ctx.handle, v.ctypes.data, len(v), t.ctypes.data, len(t),
                                   _i3(cells), C.c_float(cell_size), _f3(bmin), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def load(cls, ctx, path, bmin=None, bmax=None) -> "DistanceMap":
        """A map file of RigidBody::storeDistanceMapToFile (rigidbody.cpp:
        1318-1376), converted to binary32 on upload.  The file holds no
        bounding box: give the body's box, or None for the first and last
        lattice vertex."""
        if (bmin is None) != (bmax is None):
            raise ValueError("give both box corners or neither")
        h = C.c_void_p()
        check(lib.fc_distmap_load(ctx.handle, str(path).encode(),
                                  _f3(bmin) if bmin is not None else None,
                                  _f3(bmax) if bmax is not None else None, C.byref(h)))
        return cls(ctx, h)

    def save(self, path) -> None:
        """Write the map in the same text format (nine significant digits)."""
        check(lib.fc_distmap_save(self.ctx.handle, self.handle, str(path).encode()))

    def fbm(self) -> np.ndarray:
        """FBM state of every lattice vertex (0 when none were loaded)."""
        out = np.zeros(self.num_vertices, np.int32)
        check(lib.fc_distmap_get_fbm(self.handle, out.ctypes.data))
        return out

    def set_fbm(self, states=None) -> None:
        if states is None:
            check(lib.fc_distmap_set_fbm(self.handle, None))
            return
        a = np.ascontiguousarray(states, np.int32).reshape(-1)
        if len(a) != self.num_vertices:
            raise ValueError("one state per lattice vertex")
        check(lib.fc_distmap_set_fbm(self.handle, a.ctypes.data))

    def close(self) -> None:
        if getattr(self, "handle", None):
            check(lib.fc_distmap_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- data ---------------------------------------------------------------
    def download(self) -> dict:
        V = self.num_vertices
        out = {k: np.zeros((V, 3), np.float32) for k in ("vertex_coords", "normals",
                                                          "contact_points")}
        out["distance"] = np.zeros(V, np.float32)
        check(lib.fc_distmap_download(self.ctx.handle, self.handle,
                                      out["vertex_coords"].ctypes.data,
                                      out["normals"].ctypes.data,
                                      out["contact_points"].ctypes.data,
                                      out["distance"].ctypes.data))
        return out

    # -- queries ------------------------------------------------------------
    def query(self, points):
        """queryMap for host points (n,3): distance, closest point, normal."""
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        n = len(pts)
        dp = DeviceArray.from_host(self.ctx, pts)
        dd = DeviceArray(self.ctx, 4 * max(n, 1))
        dc = DeviceArray(self.ctx, 12 * max(n, 1))
        dn = DeviceArray(self.ctx, 12 * max(n, 1))
        check(lib.fc_distmap_query(self.ctx.handle, self.handle, dp.ptr, n, dd.ptr, dc.ptr, dn.ptr))
        self.ctx.sync()
        return (dd.download(np.float32, n), dc.download(np.float32, 3 * n).reshape(n, 3),
                dn.download(np.float32, 3 * n).reshape(n, 3))

    def query_fbm(self, points) -> np.ndarray:
        """queryFBM for host points (n,3): 1 inside the body (interpolated
        distance <= 0 within the box), else 0."""
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        n = len(pts)
        dp = DeviceArray.from_host(self.ctx, pts)
        do = DeviceArray(self.ctx, 4 * max(n, 1))
        check(lib.fc_distmap_query_fbm(self.ctx.handle, self.handle, dp.ptr, n, do.ptr))
        self.ctx.sync()
        return do.download(np.int32, n)

    def mesh_contacts(self, vertices, info: _lib.TransInfo, threshold: float = 0.0015,
                      capacity: int | None = None):
        """ColliderMeshMesh<gpu>::collide: the vertices of body 1 (model
        frame, host (n,3)) closer than `threshold` to this map's body, in
        vertex order: (contact points, normals, vertex ids)."""
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        n = len(v)
        cap = n if capacity is None else int(capacity)
        dv = DeviceArray.from_host(self.ctx, v)
        dc = DeviceArray(self.ctx, 12 * max(cap, 1))
        dn = DeviceArray(self.ctx, 12 * max(cap, 1))
        di = DeviceArray(self.ctx, 4 * max(cap, 1))
        cnt = C.c_uint32()
        check(lib.fc_mesh_mesh_contacts(self.ctx.handle, self.handle, dv.ptr, n, C.byref(info),
                                        float(threshold), dc.ptr, dn.ptr, di.ptr, cap,
                                        C.byref(cnt)))
        k = min(cnt.value, cap)
        return (dc.download(np.float32, 3 * k).reshape(k, 3),
                dn.download(np.float32, 3 * k).reshape(k, 3), di.download(np.uint32, k),
                cnt.value)

    def eval(self, points, info: _lib.TransInfo):
        """eval_distmap: returns (cps, normals, distance); cps/normals are 0
        where the point is outside the box (distance 1000)."""
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        n = len(pts)
        dp = DeviceArray.from_host(self.ctx, pts)
        zeros = np.zeros((max(n, 1), 3), np.float32)
        dc = DeviceArray.from_host(self.ctx, zeros)
        dn = DeviceArray.from_host(self.ctx, zeros)
        dd = DeviceArray(self.ctx, 4 * max(n, 1))
        check(lib.fc_distmap_eval(self.ctx.handle, self.handle, dp.ptr, n, C.byref(info), dc.ptr,
                                  dn.ptr, dd.ptr))
        self.ctx.sync()
        return (dc.download(np.float32, 3 * n).reshape(n, 3),
                dn.download(np.float32, 3 * n).reshape(n, 3), dd.download(np.float32, n))


def mesh_signed_distance(ctx, vertices, triangles, points):
    """Exact signed distance and closest point (binary64) of host points."""
    v = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    dv, dt, dp = (DeviceArray.from_host(ctx, a) for a in (v, t, p))
    n = len(p)
    ds = DeviceArray(ctx, 8 * max(n, 1))
    dc = DeviceArray(ctx, 24 * max(n, 1))
    check(lib.fc_mesh_signed_distance(ctx.handle, dv.ptr, len(v), dt.ptr, len(t), dp.ptr, n,
                                      ds.ptr, dc.ptr))
    ctx.sync()
    return ds.download(np.float64, n), dc.download(np.float64, 3 * n).reshape(n, 3)


def read_map_file(path) -> dict:
    """Parse a map file into host binary64 arrays, as the reference's reader
    fills DistanceMap<Real> (no device needed): cells, cell_size, dim,
    vertex_coords, contact_points, normals (n,3), distance (n), fbm (n)."""
    cells = (C.c_int32 * 3)()
    cs = C.c_double()
    dim = (C.c_int32 * 2)()
    p = str(path).encode()
    check(lib.fc_distmap_file_read(p, cells, C.byref(cs), dim, None, None, None, None, None))
    V = int(np.prod([c + 1 for c in cells]))
    out = {k: np.zeros((V, 3), np.float64) for k in ("vertex_coords", "contact_points", "normals")}
    out["distance"] = np.zeros(V, np.float64)
    out["fbm"] = np.zeros(V, np.int32)
    check(lib.fc_distmap_file_read(p, cells, C.byref(cs), dim, out["vertex_coords"].ctypes.data,
                                   out["contact_points"].ctypes.data, out["normals"].ctypes.data,
                                   out["distance"].ctypes.data, out["fbm"].ctypes.data))
    out.update(cells=tuple(cells), cell_size=cs.value, dim=tuple(dim))
    return out


def write_map_file(path, cells, cell_size, vertex_coords, contact_points, normals, distance,
                   fbm=None, digits: int = 6) -> None:
    """Write host arrays in the map-file format with `digits` significant
    digits (6 is what the reference's default stream formatting writes)."""
    cells = [int(c) for c in cells]
    V = int(np.prod([c + 1 for c in cells]))
    f = f"%.{digits}g"
    fbm = np.zeros(V, np.int64) if fbm is None else np.asarray(fbm).reshape(-1)
    with open(path, "w") as fh:
        fh.write(f"{cells[0] + 1} {(cells[1] + 1) * (cells[2] + 1)}\n")
        fh.write(f"{cells[0]} {cells[1]} {cells[2]}\n")
        fh.write((f % float(cell_size)) + "\n")
        for arr in (vertex_coords, contact_points, normals):
            a = np.asarray(arr, np.float64).reshape(V, 3)
            fh.writelines(f"{f % x} {f % y} {f % z}\n" for x, y, z in a)
        fh.writelines((f % d) + "\n" for d in np.asarray(distance, np.float64).reshape(V))
        fh.writelines(f"{int(s)}\n" for s in fbm)

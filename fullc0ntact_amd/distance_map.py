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

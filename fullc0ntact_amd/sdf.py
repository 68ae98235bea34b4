"""SdfSurrogate: per-body MLP signed-distance surrogates on the GPU (training
by the library's own forward/backward kernels, BF16 MFMA inference), plus
training-set generation from the exact point-mesh distance."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import DeviceArray, check, lib

FP32 = 0
BF16 = 1


class SdfSurrogate:
    def __init__(self, ctx: _lib.Context, centers, half_extents, freq_sigma: float = 1.0,
                 seed: int = 1973):
        centers = np.ascontiguousarray(centers, np.float32).reshape(-1, 3)
        half = np.ascontiguousarray(half_extents, np.float32).reshape(-1)
        self.ctx = ctx
        self.num_bodies = len(half)
        h = C.c_void_p()
        check(lib.fc_sdf_create(ctx.handle, self.num_bodies, centers.ctypes.data,
                                half.ctypes.data, C.c_float(freq_sigma), int(seed), C.byref(h)))
        self.handle = h
        ctx.register(self)
        self.nparams = int(lib.fc_sdf_num_params())
        self.stride = int(lib.fc_sdf_param_stride())

    def close(self) -> None:
        if getattr(self, "handle", None):
            check(lib.fc_sdf_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        """(params[B, stride], freqs[B, 16, 3], frames[B, 4])."""
        p = np.zeros((self.num_bodies, self.stride), np.float32)
        f = np.zeros((self.num_bodies, 16, 3), np.float32)
        fr = np.zeros((self.num_bodies, 4), np.float32)
        check(lib.fc_sdf_get_params(self.handle, p.ctypes.data, f.ctypes.data, fr.ctypes.data))
        return p, f, fr

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32).reshape(self.num_bodies, self.stride)
        check(lib.fc_sdf_set_params(self.handle, p.ctypes.data))

    def loss_grad(self, points_dev, labels_dev, samples_per_body: int):
        g = np.zeros((self.num_bodies, self.stride), np.float32)
        loss = np.zeros(self.num_bodies, np.float32)
        check(lib.fc_sdf_loss_grad(self.handle, points_dev.ptr, labels_dev.ptr,
                                   int(samples_per_body), g.ctypes.data, loss.ctypes.data))
        return g, loss

    def train(self, points_dev, labels_dev, samples_per_body: int, batch: int, epochs: int,
              lr: float = 1e-3, lr_decay: float = 1.0) -> np.ndarray:
        hist = np.zeros((epochs, self.num_bodies), np.float32)
        check(lib.fc_sdf_train(self.handle, points_dev.ptr, labels_dev.ptr, int(samples_per_body),
                               int(batch), int(epochs), C.c_float(lr), C.c_float(lr_decay),
                               hist.ctypes.data))
        return hist

    def map_margin(self, net: int, distance_map, radius: float, band: float) -> dict:
        """Statistics of network `net` (BF16 inference) against a distance
        map of its shape over the lattice vertices within radius + band of
        the surface: the largest y - radius, the largest secant slope
        between x-neighbours, the vertex count and the largest |y - d|."""
        out = np.zeros(4, np.float32)
        check(lib.fc_sdf_map_margin(self.handle, int(net), distance_map.handle,
                                    C.c_float(radius), C.c_float(band), out.ctypes.data))
        return {"over": float(out[0]), "slope": float(out[1]), "vertices": int(out[2]),
                "max_err": float(out[3])}

    def eval(self, points, bodies=None, precision: int = BF16) -> np.ndarray:
        """Signed distances of host model-frame points (n, 3)."""
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        n = len(pts)
        dp = DeviceArray.from_host(self.ctx, pts)
        db = DeviceArray.from_host(self.ctx, np.ascontiguousarray(bodies, np.uint32)) \
            if bodies is not None else None
        dout = DeviceArray(self.ctx, 4 * max(n, 1))
        check(lib.fc_sdf_eval(self.handle, db.ptr if db else None, dp.ptr, n, dout.ptr,
                              int(precision)))
        self.ctx.sync()
        return dout.download(np.float32, n)


def training_samples(vertices, triangles, n: int, center, half_extent: float,
                     surface_frac: float = 0.5, sigma: float = 0.05, seed: int = 1973):
    """n model-frame sample points for one body: surface_frac of them on
    the mesh surface (area-weighted) displaced by N(0, (sigma*halfExtent)^2),
    the rest uniform in the normalisation cube center +- halfExtent."""
    rng = np.random.default_rng(seed)
    v = np.asarray(vertices, np.float64)
    f = np.asarray(triangles, np.int64)
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)
    ns = int(n * surface_frac)
    tri = rng.choice(len(f), ns, p=area / area.sum())
    r1, r2 = rng.random(ns), rng.random(ns)
    s = np.sqrt(r1)
    surf = (1 - s)[:, None] * a[tri] + (s * (1 - r2))[:, None] * b[tri] + (s * r2)[:, None] * c[tri]
    surf += rng.normal(0.0, sigma * half_extent, size=surf.shape)
    uni = np.asarray(center) + rng.uniform(-half_extent, half_extent, size=(n - ns, 3))
    pts = np.concatenate([surf, uni])
    # shuffled (the trainer takes minibatches of consecutive samples) by a
    # permutation index: an in-place row shuffle is ~20x slower at 2^20 rows
    return pts[rng.permutation(len(pts))]

"""ctypes binding of liboracle.so, the CPU restatement of the particle step.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the CPU baseline.  The product
(fullc0ntact_amd/) never imports it.

Parity status of each function (what pins it, see tests/test_oracle_*.py):
  hash          known-answer cell coordinates incl. faces, wrap, G=64/256
  sort          numpy lexsort on (hash, index); stability
  reorder       closed-form cell-table invariants
  pair force    two-sphere closed forms (head-on, tangential, separation)
  collide       brute-force all-pairs on small boxes; momentum invariant;
                binary64 twin agreement
  integrate     ballistic closed form and wall reflection
  reset         lattice geometry (spacing, bounds, jitter range)
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


class OrcParams(C.Structure):
    _fields_ = [
        ("colliderPos", C.c_float * 3),
        ("colliderRadius", C.c_float),
        ("gravity", C.c_float * 3),
        ("globalDamping", C.c_float),
        ("particleRadius", C.c_float),
        ("gridSize", C.c_uint32 * 3),
        ("numCells", C.c_uint32),
        ("worldOrigin", C.c_float * 3),
        ("cellSize", C.c_float * 3),
        ("numBodies", C.c_uint32),
        ("maxParticlesPerCell", C.c_uint32),
        ("spring", C.c_float),
        ("damping", C.c_float),
        ("shear", C.c_float),
        ("attraction", C.c_float),
        ("boundaryDamping", C.c_float),
    ]


class OrcDistMap(C.Structure):
    _fields_ = [("cells", C.c_int * 3), ("cellSize", C.c_float), ("bmin", C.c_float * 3),
                ("bmax", C.c_float * 3), ("xs", C.c_void_p), ("ys", C.c_void_p),
                ("zs", C.c_void_p), ("dist", C.c_void_p), ("cp", C.c_void_p),
                ("nrm", C.c_void_p)]


class OrcTransInfo(C.Structure):
    _fields_ = [("origin0", C.c_float * 3), ("origin1", C.c_float * 3), ("m2w0", C.c_float * 9),
                ("w2m0", C.c_float * 9), ("m2w1", C.c_float * 9)]


class OrcState(C.Structure):
    _fields_ = [("pos4", C.c_void_p), ("vel4", C.c_void_p), ("sortedPos4", C.c_void_p),
                ("sortedVel4", C.c_void_p), ("hash", C.c_void_p), ("index", C.c_void_p),
                ("cellStart", C.c_void_p), ("cellEnd", C.c_void_p), ("n", C.c_uint32)]


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        build()
    L = C.CDLL(str(LIB_PATH))
    vp, u32, f = C.c_void_p, C.c_uint32, C.c_float
    P = C.POINTER(OrcParams)
    sig = {
        "orc_params_size": (C.c_int, []),
        "orc_set_threads": (None, [C.c_int]),
        "orc_get_threads": (C.c_int, []),
        "orc_integrate": (None, [vp, vp, u32, f, P]),
        "orc_hash_point": (u32, [f, f, f, P]),
        "orc_calc_hash": (None, [vp, vp, vp, u32, P]),
        "orc_sort_pairs": (None, [vp, vp, u32]),
        "orc_reorder": (None, [vp, vp, vp, vp, vp, vp, vp, vp, u32, u32]),
        "orc_collide": (None, [vp, vp, vp, vp, vp, vp, u32, P]),
        "orc_eval_forces": (None, [vp, vp, vp, vp, vp, vp, vp, u32, P]),
        "orc_collide_f64": (None, [vp, vp, vp, vp, vp, vp, u32, P]),
        "orc_pair_force": (None, [vp, vp, vp, vp, f, f, f, P, vp]),
        "orc_update": (None, [C.POINTER(OrcState), f, P]),
        "orc_run": (None, [C.POINTER(OrcState), f, u32, P]),
        "orc_reset": (None, [vp, vp, u32, C.c_int, C.c_uint64, f, f, f]),
        "orc_energy_momentum": (None, [vp, u32, vp]),
        "orc_distmap_query_batch": (None, [C.POINTER(OrcDistMap), vp, u32, vp, vp, vp]),
        "orc_integrate_distmap": (None, [vp, vp, u32, f, P, C.POINTER(OrcDistMap)]),
        "orc_distmap_eval": (None, [C.POINTER(OrcDistMap), vp, u32, C.POINTER(OrcTransInfo), vp,
                                    vp, vp]),
        "orc_point_triangle_dist2": (C.c_double, [vp, vp, vp, vp, vp]),
        "orc_mesh_sdf": (None, [vp, vp, u32, vp, u32, vp, vp, vp]),
        "orc_update_distmap": (None, [C.POINTER(OrcState), f, P, C.POINTER(OrcDistMap)]),
        "orc_integrate_poly": (None, [vp, vp, u32, f, P]),
        "orc_collide_poly": (None, [vp, vp, vp, vp, vp, vp, u32, P]),
        "orc_contact_counts": (None, [vp, vp, vp, vp, u32, P, C.c_int]),
        "orc_integrate_distmap_ex": (None, [vp, vp, u32, f, P, C.POINTER(OrcDistMap), C.c_int]),
        "orc_update_ex": (None, [C.POINTER(OrcState), f, P, C.POINTER(OrcDistMap), C.c_int]),
        "orc_sdf_num_params": (C.c_int, []),
        "orc_sdf_forward": (None, [vp, vp, vp, vp, u32, vp]),
        "orc_sdf_loss_grad": (None, [vp, vp, vp, vp, vp, u32, vp, vp]),
        "orc_sdf_forward_bf16": (None, [vp, vp, vp, vp, u32, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    assert L.orc_params_size() == 104
    return L


lib = _load()


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def params_from(sim) -> OrcParams:
    """Copy a fullc0ntact_amd SimParams (same 104-byte layout) or a dict."""
    o = OrcParams()
    if isinstance(sim, dict):
        for k, v in sim.items():
            if isinstance(v, (tuple, list)):
                arr = getattr(o, k)
                for i, x in enumerate(v):
                    arr[i] = x
            else:
                setattr(o, k, v)
        return o
    C.memmove(C.byref(o), C.byref(sim), 104)
    return o


def make_params(n: int, grid: int, radius: float, **overrides) -> OrcParams:
    """ParticleSystem defaults (particleSystem.cpp:48-68) for a cubic grid."""
    c = float(np.float32(radius) * np.float32(2.0))
    d = dict(colliderPos=(-1.2, -0.8, 0.8), colliderRadius=0.2, gravity=(0.0, 0.0, -0.0003),
             globalDamping=1.0, particleRadius=radius, gridSize=(grid, grid, grid),
             numCells=grid ** 3, worldOrigin=(-1.0, -1.0, -1.0), cellSize=(c, c, c),
             numBodies=n, maxParticlesPerCell=0, spring=0.5, damping=0.02, shear=0.1,
             attraction=0.0, boundaryDamping=-0.5)
    d.update(overrides)
    return params_from(d)


def set_threads(t: int) -> None:
    lib.orc_set_threads(int(t))


def integrate(pos4, vel4, dt, P):
    pos4 = np.array(pos4, np.float32, copy=True, order="C")
    vel4 = np.array(vel4, np.float32, copy=True, order="C")
    lib.orc_integrate(_p(pos4), _p(vel4), len(pos4), dt, C.byref(P))
    return pos4, vel4


def calc_hash(pos4, P):
    pos4 = np.ascontiguousarray(pos4, np.float32)
    n = len(pos4)
    h = np.empty(n, np.uint32)
    i = np.empty(n, np.uint32)
    lib.orc_calc_hash(_p(h), _p(i), _p(pos4), n, C.byref(P))
    return h, i


def hash_point(x, y, z, P) -> int:
    return int(lib.orc_hash_point(x, y, z, C.byref(P)))


def sort_pairs(h, i):
    h = np.array(h, np.uint32, copy=True)
    i = np.array(i, np.uint32, copy=True)
    lib.orc_sort_pairs(_p(h), _p(i), len(h))
    return h, i


def reorder(h, i, pos4, vel4, numCells):
    n = len(h)
    cs = np.empty(numCells, np.uint32)
    ce = np.empty(numCells, np.uint32)
    sp = np.empty((n, 4), np.float32)
    sv = np.empty((n, 4), np.float32)
    h = np.ascontiguousarray(h, np.uint32)
    i = np.ascontiguousarray(i, np.uint32)
    pos4 = np.ascontiguousarray(pos4, np.float32)
    vel4 = np.ascontiguousarray(vel4, np.float32)
    lib.orc_reorder(_p(cs), _p(ce), _p(sp), _p(sv), _p(h), _p(i), _p(pos4), _p(vel4), n, numCells)
    return cs, ce, sp, sv


def collide(sp, sv, index, cs, ce, P, newVel4=None):
    n = len(sp)
    out = np.zeros((n, 4), np.float32) if newVel4 is None else np.array(newVel4, np.float32)
    sp = np.ascontiguousarray(sp, np.float32)
    sv = np.ascontiguousarray(sv, np.float32)
    index = np.ascontiguousarray(index, np.uint32)
    lib.orc_collide(_p(out), _p(sp), _p(sv), _p(index), _p(np.ascontiguousarray(cs)),
                    _p(np.ascontiguousarray(ce)), n, C.byref(P))
    return out


def eval_forces(sp, sv, index, cs, ce, P, newVel4=None):
    """evalForces_launcher (particledem_um.cu:538-589): (newVel4, force3),
    the 27-cell force without the cursor sphere, both at the original
    indices."""
    n = len(sp)
    out = np.zeros((n, 4), np.float32) if newVel4 is None else np.array(newVel4, np.float32)
    force = np.zeros((n, 3), np.float32)
    sp = np.ascontiguousarray(sp, np.float32)
    sv = np.ascontiguousarray(sv, np.float32)
    index = np.ascontiguousarray(index, np.uint32)
    lib.orc_eval_forces(_p(out), _p(force), _p(sp), _p(sv), _p(index),
                        _p(np.ascontiguousarray(cs)), _p(np.ascontiguousarray(ce)), n, C.byref(P))
    return out, force


def integrate_poly(pos4, vel4, dt, P):
    pos4 = np.array(pos4, np.float32, copy=True, order="C")
    vel4 = np.array(vel4, np.float32, copy=True, order="C")
    lib.orc_integrate_poly(_p(pos4), _p(vel4), len(pos4), dt, C.byref(P))
    return pos4, vel4


def collide_poly(sp, sv, index, cs, ce, P, newVel4=None):
    n = len(sp)
    out = np.zeros((n, 4), np.float32) if newVel4 is None else np.array(newVel4, np.float32)
    sp = np.ascontiguousarray(sp, np.float32)
    sv = np.ascontiguousarray(sv, np.float32)
    index = np.ascontiguousarray(index, np.uint32)
    lib.orc_collide_poly(_p(out), _p(sp), _p(sv), _p(index), _p(np.ascontiguousarray(cs)),
                         _p(np.ascontiguousarray(ce)), n, C.byref(P))
    return out


def contact_counts(sp, cs, ce, P, poly=False):
    """Contacts per sorted sphere (the collide's contact test)."""
    sp = np.ascontiguousarray(sp, np.float32)
    out = np.zeros(len(sp), np.uint32)
    lib.orc_contact_counts(_p(out), _p(sp), _p(np.ascontiguousarray(cs, np.uint32)),
                           _p(np.ascontiguousarray(ce, np.uint32)), len(sp), C.byref(P),
                           1 if poly else 0)
    return out


def contact_histogram(sp, cs, ce, P, nbins, poly=False):
    c = contact_counts(sp, cs, ce, P, poly)
    return np.bincount(np.minimum(c, nbins - 1), minlength=nbins).astype(np.uint32)


def collide_f64(sp, sv, index, cs, ce, P):
    n = len(sp)
    out = np.zeros((n, 3), np.float64)
    sp = np.ascontiguousarray(sp, np.float32)
    sv = np.ascontiguousarray(sv, np.float32)
    index = np.ascontiguousarray(index, np.uint32)
    lib.orc_collide_f64(_p(out), _p(sp), _p(sv), _p(index), _p(np.ascontiguousarray(cs)),
                        _p(np.ascontiguousarray(ce)), n, C.byref(P))
    return out


def pair_force(pA, pB, vA, vB, rA, rB, attraction, P):
    arrs = [np.ascontiguousarray(a, np.float32) for a in (pA, pB, vA, vB)]
    out = np.zeros(3, np.float32)
    lib.orc_pair_force(*[_p(a) for a in arrs], rA, rB, attraction, C.byref(P), _p(out))
    return out


class State:
    """Host arrays of one particle system, stepped by orc_update."""

    def __init__(self, pos4, vel4, numCells):
        n = len(pos4)
        self.pos4 = np.array(pos4, np.float32, copy=True, order="C").reshape(n, 4)
        self.vel4 = np.array(vel4, np.float32, copy=True, order="C").reshape(n, 4)
        self.sortedPos4 = np.zeros((n, 4), np.float32)
        self.sortedVel4 = np.zeros((n, 4), np.float32)
        self.hash = np.zeros(n, np.uint32)
        self.index = np.zeros(n, np.uint32)
        self.cellStart = np.full(numCells, 0xFFFFFFFF, np.uint32)
        self.cellEnd = np.full(numCells, 0xFFFFFFFF, np.uint32)
        self.n = n
        self._s = OrcState(_p(self.pos4), _p(self.vel4), _p(self.sortedPos4),
                           _p(self.sortedVel4), _p(self.hash), _p(self.index),
                           _p(self.cellStart), _p(self.cellEnd), n)

    def update(self, dt, P):
        lib.orc_update(C.byref(self._s), dt, C.byref(P))

    def run(self, dt, steps, P):
        lib.orc_run(C.byref(self._s), dt, steps, C.byref(P))


def reset(n, config, seed, radius, jitter_frac, vel_amp):
    pos = np.zeros((n, 4), np.float32)
    vel = np.zeros((n, 4), np.float32)
    lib.orc_reset(_p(pos), _p(vel), n, int(config), int(seed), radius, jitter_frac, vel_amp)
    return pos, vel


def energy_momentum(vel4) -> np.ndarray:
    vel4 = np.ascontiguousarray(vel4, np.float32)
    out = np.zeros(4, np.float64)
    lib.orc_energy_momentum(_p(vel4), len(vel4), _p(out))
    return out


# ---------------------------------------------------------------- distance map
class DistMap:
    """Host arrays of a tensor-product distance map for the oracle."""

    def __init__(self, cells, cell_size, bmin, bmax, xs, ys, zs, dist, cp, nrm):
        self.arrays = [np.ascontiguousarray(a, np.float32) for a in (xs, ys, zs, dist, cp, nrm)]
        self.s = OrcDistMap()
        for a in range(3):
            self.s.cells[a] = int(cells[a])
            self.s.bmin[a] = float(bmin[a])
            self.s.bmax[a] = float(bmax[a])
        self.s.cellSize = float(cell_size)
        self.s.xs, self.s.ys, self.s.zs, self.s.dist, self.s.cp, self.s.nrm = [
            _p(a) for a in self.arrays]

    @classmethod
    def from_reference_layout(cls, cells, cell_size, bmin, bmax, coords, nrm, cp, dist):
        cx, cy, cz = (int(c) for c in cells)
        coords = np.asarray(coords, np.float32).reshape(-1, 3)
        ix, isl = cx + 1, (cx + 1) * (cy + 1)
        xs = coords[np.arange(cx + 1), 0]
        ys = coords[np.arange(cy + 1) * ix, 1]
        zs = coords[np.arange(cz + 1) * isl, 2]
        return cls(cells, cell_size, bmin, bmax, xs, ys, zs, dist, np.asarray(cp).reshape(-1),
                   np.asarray(nrm).reshape(-1))


def distmap_query(m: DistMap, pts):
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    n = len(pts)
    d = np.zeros(n, np.float32)
    cp = np.zeros((n, 3), np.float32)
    nr = np.zeros((n, 3), np.float32)
    lib.orc_distmap_query_batch(C.byref(m.s), _p(pts), n, _p(d), _p(cp), _p(nr))
    return d, cp, nr


def integrate_distmap(pos4, vel4, dt, P, m: DistMap):
    pos4 = np.array(pos4, np.float32, copy=True, order="C")
    vel4 = np.array(vel4, np.float32, copy=True, order="C")
    lib.orc_integrate_distmap(_p(pos4), _p(vel4), len(pos4), dt, C.byref(P), C.byref(m.s))
    return pos4, vel4


def distmap_eval(m: DistMap, v, info: OrcTransInfo, cps=None, normals=None):
    v = np.ascontiguousarray(v, np.float32).reshape(-1, 3)
    n = len(v)
    cps = np.zeros((n, 3), np.float32) if cps is None else np.array(cps, np.float32)
    normals = np.zeros((n, 3), np.float32) if normals is None else np.array(normals, np.float32)
    dist = np.zeros(n, np.float32)
    lib.orc_distmap_eval(C.byref(m.s), _p(v), n, C.byref(info), _p(cps), _p(normals), _p(dist))
    return cps, normals, dist


def point_triangle_dist2(q, a, b, c):
    arrs = [np.ascontiguousarray(x, np.float64) for x in (q, a, b, c)]
    cp = np.zeros(3)
    d2 = lib.orc_point_triangle_dist2(*[_p(x) for x in arrs], _p(cp))
    return d2, cp


def mesh_sdf(verts, tris, pts):
    verts = np.ascontiguousarray(verts, np.float64)
    tris = np.ascontiguousarray(tris, np.uint32)
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    n = len(pts)
    sdf = np.zeros(n)
    cp = np.zeros((n, 3))
    w = np.zeros(n)
    lib.orc_mesh_sdf(_p(verts), _p(tris), len(tris), _p(pts), n, _p(sdf), _p(cp), _p(w))
    return sdf, cp, w


def update_distmap(state: "State", dt, P, m: DistMap):
    lib.orc_update_distmap(C.byref(state._s), dt, C.byref(P), C.byref(m.s))


def update_ex(state: "State", dt, P, m: "DistMap | None" = None, poly: bool = False):
    """One update with a radius mode and an optional distance-map boundary."""
    lib.orc_update_ex(C.byref(state._s), dt, C.byref(P), C.byref(m.s) if m is not None else None,
                      1 if poly else 0)


def integrate_distmap_poly(pos4, vel4, dt, P, m: DistMap):
    pos4 = np.array(pos4, np.float32, copy=True, order="C")
    vel4 = np.array(vel4, np.float32, copy=True, order="C")
    lib.orc_integrate_distmap_ex(_p(pos4), _p(vel4), len(pos4), dt, C.byref(P), C.byref(m.s), 1)
    return pos4, vel4


# ------------------------------------------------- the reference's own build
REF_LIB = HERE / "_ref" / "libref_distance.so"


def reference_point_triangle():
    """FullC0ntact's CDistancePointTriangle<double>::ComputeDistance, built
    from the reference sources by oracle/ref.mk; None when not built."""
    if not REF_LIB.exists():
        return None
    L = C.CDLL(str(REF_LIB))
    fn = L.ref_point_triangle_distance
    fn.restype = C.c_double
    fn.argtypes = [C.c_void_p] * 5

    def call(q, a, b, c):
        arrs = [np.ascontiguousarray(x, np.float64) for x in (q, a, b, c)]
        cp = np.zeros(3)
        d = fn(*[_p(x) for x in arrs], _p(cp))
        return d, cp
    return call


# ------------------------------------------------- signed-distance surrogate
SDF_NPARAMS = 10497


def _sdf_args(params, freq, frame, pts):
    params = np.ascontiguousarray(params, np.float32).reshape(-1)[:SDF_NPARAMS]
    freq = np.ascontiguousarray(freq, np.float32).reshape(48)
    frame = np.ascontiguousarray(frame, np.float32).reshape(4)
    pts = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
    return params, freq, frame, pts


def sdf_forward(params, freq, frame, pts, bf16=False):
    """Signed distance of model-frame points by the MLP of one body
    (binary64; bf16=True rounds weights and activations as the MFMA kernel)."""
    params, freq, frame, pts = _sdf_args(params, freq, frame, pts)
    out = np.zeros(len(pts))
    fn = lib.orc_sdf_forward_bf16 if bf16 else lib.orc_sdf_forward
    fn(_p(params), _p(freq), _p(frame), _p(pts), len(pts), _p(out))
    return out


def sdf_loss_grad(params, freq, frame, pts, labels):
    params, freq, frame, pts = _sdf_args(params, freq, frame, pts)
    labels = np.ascontiguousarray(labels, np.float32).reshape(-1)
    grad = np.zeros(SDF_NPARAMS)
    loss = np.zeros(1)
    lib.orc_sdf_loss_grad(_p(params), _p(freq), _p(frame), _p(pts), _p(labels), len(pts),
                          _p(grad), _p(loss))
    return grad, float(loss[0])

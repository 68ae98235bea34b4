"""Rigid bodies moved by the particle step, their FEATFLOW entry points and
the GPU mesh-mesh contacts, against the oracle.

 - coupling: one sphere per body with initParticleSystem's parameters;
   each step is 10 updates of (float)dt / 10 and hands the spheres' state
   to the bodies in binary64 -- bit for bit the oracle's 10 updates;
   setting a body moves its sphere;
 - the Fortran entry points (called as Fortran does, every argument by
   reference) read and write the bound coupling, and startcollisionpipeline_
   equals a direct coupling step with the time step of settimestep_;
 - ColliderMeshMesh<gpu>::collide: the contacts are exactly the vertices
   whose oracle eval_distmap distance is below the threshold, in vertex
   order, with the oracle's contact points and normals bit for bit.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import distmap_cases as dc
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GRID = 16


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _bodies(seed=2):
    """A jittered 8^3 lattice of centres of mass (radius 1/16 spheres)."""
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(*[np.arange(8)] * 3, indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
    com = -1.0 + 1.0 / 16 + g * (2.0 / 16) * 1.01 + rng.uniform(-0.003, 0.003, (len(g), 3))
    return com.astype(np.float64)


def _oracle_state(com):
    n = len(com)
    P = orc.make_params(n, GRID, 1.0 / 16, globalDamping=1.0, gravity=(0.0, 0.0, -9.81),
                        spring=2.75, damping=0.02, shear=0.1, attraction=0.0)
    pos = np.ones((n, 4), np.float32)
    pos[:, :3] = com.astype(np.float32)
    return P, orc.State(pos, np.zeros((n, 4), np.float32), P.numCells)


def test_coupling_steps_match_oracle(gpu_ctx):
    from fullc0ntact_amd.coupling import Coupling
    com = _bodies()
    cp = Coupling(com, (GRID,) * 3, context=gpu_ctx)
    P, st = _oracle_state(com)
    dt = 0.01
    sub = np.float32(np.float32(dt) / np.float32(10))
    for step in range(3):
        cp.step(dt, 10)
        for _ in range(10):
            st.update(float(sub), P)
        got_p = np.array([cp.body(i)[0] for i in range(len(com))])
        got_v = np.array([cp.body(i)[1] for i in range(len(com))])
        assert np.array_equal(got_p, st.pos4[:, :3].astype(np.float64)), step
        assert np.array_equal(got_v, st.vel4[:, :3].astype(np.float64)), step
    assert np.any(got_v != 0.0)
    # moving a body moves its sphere
    cp.set_body(5, pos=(0.1, 0.2, -0.3), vel=(0.0, 0.01, 0.0))
    st.pos4[5, :3] = (0.1, 0.2, -0.3)
    st.vel4[5, :3] = (0.0, 0.01, 0.0)
    cp.step(dt, 10)
    for _ in range(10):
        st.update(float(sub), P)
    got_p = np.array([cp.body(i)[0] for i in range(len(com))])
    assert np.array_equal(got_p, st.pos4[:, :3].astype(np.float64))
    cp.close()


def test_fortran_entry_points(gpu_ctx):
    from fullc0ntact_amd import _lib
    from fullc0ntact_amd.coupling import Coupling
    lib = _lib.lib
    com = _bodies(7)
    a = Coupling(com, (GRID,) * 3, context=gpu_ctx)
    b = Coupling(com, (GRID,) * 3, context=gpu_ctx)
    a.bind()
    n = C.c_int(0)
    lib.getnumparticles_(C.byref(n))
    assert n.value == len(com)
    dt = C.c_double(0.02)
    lib.settimestep_(C.byref(dt))
    lib.startcollisionpipeline_()
    b.step(0.02, 10)
    d = [C.c_double() for _ in range(12)]
    for i in (0, 17, len(com) - 1):
        iid = C.c_int(i)
        lib.update_particle_state_(*[C.byref(x) for x in d], C.byref(iid))
        p, v = b.body(i)
        assert [x.value for x in d[:3]] == list(p) and [x.value for x in d[3:6]] == list(v)
        assert all(x.value == 0.0 for x in d[6:])
        lib.getpos_(C.byref(d[0]), C.byref(d[1]), C.byref(d[2]), C.byref(iid))
        assert [x.value for x in d[:3]] == list(p)
        lib.getvel_(C.byref(d[0]), C.byref(d[1]), C.byref(d[2]), C.byref(iid))
        assert [x.value for x in d[:3]] == list(v)
    r = C.c_double()
    lib.getradius_(C.byref(r), C.byref(C.c_int(3)))
    assert r.value == np.float32(1.0 / 16)
    x, y, z = C.c_double(0.25), C.c_double(-0.5), C.c_double(0.125)
    lib.setpositionid_(C.byref(x), C.byref(y), C.byref(z), C.byref(C.c_int(9)))
    lib.setvelocityid_(C.byref(z), C.byref(y), C.byref(x), C.byref(C.c_int(9)))
    p, v = a.body(9)
    assert list(p) == [0.25, -0.5, 0.125] and list(v) == [0.125, -0.5, 0.25]
    # a bad id leaves the outputs alone and records the error
    d[0].value = 42.0
    lib.getpos_(C.byref(d[0]), C.byref(d[1]), C.byref(d[2]), C.byref(C.c_int(len(com))))
    assert d[0].value == 42.0 and "out of range" in _lib.last_error()
    lib.fc_featflow_bind(None)
    a.close()
    b.close()


def test_mesh_mesh_contacts_match_oracle(gpu_ctx):
    from fullc0ntact_amd import _lib
    from fullc0ntact_amd.distance_map import DistanceMap
    v0, f0 = dc.small_sphere(0.5, 24, 12)
    gm = DistanceMap.from_mesh(gpu_ctx, v0, f0, (32, 32, 32), 1.5 / 32, (-0.75,) * 3)
    lay = gm.download()
    m = orc.DistMap.from_reference_layout(gm.cells, np.float32(gm.cell_size),
                                          np.asarray(gm.bmin, np.float32),
                                          np.asarray(gm.bmax, np.float32), lay["vertex_coords"],
                                          lay["normals"], lay["contact_points"], lay["distance"])
    # body 1: a finely tessellated sphere overlapping body 0
    v1, _ = dc.small_sphere(0.3, 96, 48)
    v1 = np.asarray(v1, np.float32)
    gi, oi = _lib.TransInfo(), orc.OrcTransInfo()
    o0, o1 = (0.0, 0.0, 0.0), (0.62, 0.05, -0.03)
    gi.origin0, gi.origin1 = _lib.Float3(*o0), _lib.Float3(*o1)
    for a in range(3):
        oi.origin0[a], oi.origin1[a] = o0[a], o1[a]
    c, s = np.cos(0.3), np.sin(0.3)
    R1 = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]], np.float32)
    for t in (gi, oi):
        for k in range(9):
            t.m2w0[k] = 1.0 if k in (0, 4, 8) else 0.0
            t.w2m0[k] = 1.0 if k in (0, 4, 8) else 0.0
            t.m2w1[k] = float(R1.ravel()[k])
    for thr in (0.0015, 0.05):
        cps, nrm, ids, count = gm.mesh_contacts(v1, gi, thr)
        rc, rn, rd = orc.distmap_eval(m, v1, oi)
        want = np.nonzero(rd < np.float32(thr))[0]
        print(f"threshold {thr}: {count} contacts of {len(v1)} vertices", flush=True)
        assert count == len(want) and np.array_equal(ids, want.astype(np.uint32))
        assert np.array_equal(bits(cps), bits(rc[want])) and np.array_equal(bits(nrm), bits(rn[want]))
    assert count > 0
    # a short buffer gets the first contacts and the full count
    cps2, _, ids2, count2 = gm.mesh_contacts(v1, gi, 0.05, capacity=5)
    assert count2 == count and np.array_equal(ids2, ids[:5])
    gm.close()

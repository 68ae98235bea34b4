"""Pins the CPU oracle (oracle/fc_oracle.c) before it is trusted as the
parity checker.  Each check is a closed form, a known answer, an invariant
or a brute-force restatement chosen so that a plausible mistake (an axis
transposed in the hash, a sign flip in the spring, a dropped shear term, a
wrong wall) fails at least one of them.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc

EMPTY = 0xFFFFFFFF


# --------------------------------------------------------------------------
# calcGridPos / calcGridHash (particles_kernel.cu:78-94)
# --------------------------------------------------------------------------
def test_hash_known_answers_grid64():
    P = orc.make_params(4096, 64, 1.0 / 64.0)  # cell = 1/32, origin -1
    assert orc.hash_point(-1 + 1 / 64, -1 + 1 / 64, -1 + 1 / 64, P) == 0
    # x exactly on the face of cell 5, y inside cell 3, z inside cell 7:
    # hash = (z*Gy + y)*Gx + x
    x = np.float32(-1.0 + 5.0 / 32.0)
    y = np.float32(-1.0 + 3.0 / 32.0 + 1.0 / 64.0)
    z = np.float32(-1.0 + 7.0 / 32.0 + 0.001)
    assert orc.hash_point(x, y, z, P) == (7 * 64 + 3) * 64 + 5
    # below the origin: floor gives -1, which wraps to G-1
    assert orc.hash_point(-1.001, -1 + 1 / 64, -1 + 1 / 64, P) == 63
    assert orc.hash_point(-1 + 1 / 64, -1.001, -1 + 1 / 64, P) == 63 * 64
    assert orc.hash_point(-1 + 1 / 64, -1 + 1 / 64, -1.001, P) == 63 * 64 * 64
    # x = +1 is cell 64, which wraps to 0; x = 3 is cell 128 -> 0
    assert orc.hash_point(1.0, -0.99, -0.99, P) == 0
    assert orc.hash_point(3.0, -0.99, -0.99, P) == 0


def test_hash_non_cubic_grid():
    P = orc.make_params(1, 8, 0.125, gridSize=(8, 4, 2), numCells=64, cellSize=(0.25, 0.25, 0.25))
    # cell (5, 3, 1): (1*4 + 3)*8 + 5 = 61
    assert orc.hash_point(-1 + 5.5 * 0.25, -1 + 3.5 * 0.25, -1 + 1.5 * 0.25, P) == 61
    # y = 3.5 cells but Gy = 4 -> 3; y = 4.5 cells wraps to 0
    assert orc.hash_point(-1 + 0.5 * 0.25, -1 + 4.5 * 0.25, -1 + 0.5 * 0.25, P) == 0


def test_hash_grid256_upper_corner():
    P = orc.make_params(1 << 20, 256, 1.0 / 128.0)  # cell 1/64 covers [-1, 3)
    assert orc.hash_point(2.999, 2.999, 2.999, P) == 256 ** 3 - 1
    assert orc.hash_point(-0.999, -0.999, -0.999, P) == 0


def test_hash_non_finite_is_deterministic():
    P = orc.make_params(1, 64, 1.0 / 64.0)
    c = -1 + 1 / 64
    assert orc.hash_point(np.nan, c, c, P) == 0
    assert orc.hash_point(np.inf, c, c, P) == 63      # saturates to 2^31-1, & 63
    assert orc.hash_point(-np.inf, c, c, P) == 0      # saturates to -2^31, & 63


def test_hash_matches_float32_closed_form_random():
    rng = np.random.default_rng(5)
    P = orc.make_params(5000, 64, 1.0 / 64.0)
    pos = np.zeros((5000, 4), np.float32)
    pos[:, :3] = rng.uniform(-1.3, 1.3, size=(5000, 3)).astype(np.float32)
    h, idx = orc.calc_hash(pos, P)
    cell = np.float32(1.0 / 32.0)
    g = np.floor((pos[:, :3] - np.float32(-1.0)) / cell).astype(np.int64) & 63
    want = (g[:, 2] * 64 + g[:, 1]) * 64 + g[:, 0]
    assert np.array_equal(h, want.astype(np.uint32))
    assert np.array_equal(idx, np.arange(5000, dtype=np.uint32))


# --------------------------------------------------------------------------
# sort_by_key (particleSystem.cu:265-270)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("n,maxkey", [(1, 10), (2, 1), (1000, 7), (100_000, 1 << 24), (70_000, 1 << 31)])
def test_sort_pairs_equals_lexsort(n, maxkey):
    rng = np.random.default_rng(n)
    h = rng.integers(0, maxkey, size=n, dtype=np.int64).astype(np.uint32)
    i = np.arange(n, dtype=np.uint32)
    sh, si = orc.sort_pairs(h, i)
    order = np.lexsort((i, h))
    assert np.array_equal(sh, h[order])
    assert np.array_equal(si, i[order])


def test_sort_is_stable_for_non_identity_values():
    h = np.array([3, 1, 3, 1, 2, 3], np.uint32)
    v = np.array([50, 40, 30, 20, 10, 0], np.uint32)
    sh, sv = orc.sort_pairs(h, v)
    assert sh.tolist() == [1, 1, 2, 3, 3, 3]
    assert sv.tolist() == [40, 20, 10, 50, 30, 0]


def test_sort_threads_do_not_change_result():
    rng = np.random.default_rng(1)
    h = rng.integers(0, 1 << 18, size=200_000).astype(np.uint32)
    i = np.arange(200_000, dtype=np.uint32)
    orc.set_threads(1)
    a = orc.sort_pairs(h, i)
    orc.set_threads(4)
    b = orc.sort_pairs(h, i)
    orc.set_threads(1)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# --------------------------------------------------------------------------
# reorderDataAndFindCellStart (particles_kernel.cu:119-182)
# --------------------------------------------------------------------------
def test_reorder_small_known_answer():
    h = np.array([2, 2, 5, 7, 7, 7], np.uint32)
    idx = np.array([4, 0, 5, 1, 3, 2], np.uint32)
    pos = np.arange(24, dtype=np.float32).reshape(6, 4)
    vel = -pos
    cs, ce, sp, sv = orc.reorder(h, idx, pos, vel, 8)
    want_cs = [EMPTY, EMPTY, 0, EMPTY, EMPTY, 2, EMPTY, 3]
    want_ce = [EMPTY, EMPTY, 2, EMPTY, EMPTY, 3, EMPTY, 6]
    assert cs.tolist() == want_cs and ce.tolist() == want_ce
    assert np.array_equal(sp, pos[idx]) and np.array_equal(sv, vel[idx])


def cell_table_invariants(h_sorted, cs, ce, n):
    occ = np.unique(h_sorted)
    assert np.all(cs[occ] != EMPTY) and np.all(ce[occ] != EMPTY)
    empty = np.ones(len(cs), bool)
    empty[occ] = False
    assert np.all(cs[empty] == EMPTY) and np.all(ce[empty] == EMPTY)
    assert int(np.sum(ce[occ].astype(np.int64) - cs[occ])) == n
    k = np.arange(n)
    assert np.all((cs[h_sorted] <= k) & (k < ce[h_sorted]))


def test_reorder_invariants_random():
    rng = np.random.default_rng(2)
    n, C = 20_000, 4096
    h = np.sort(rng.integers(0, C, size=n).astype(np.uint32))
    idx = rng.permutation(n).astype(np.uint32)
    pos = rng.standard_normal((n, 4)).astype(np.float32)
    cs, ce, sp, sv = orc.reorder(h, idx, pos, pos, C)
    cell_table_invariants(h, cs, ce, n)
    assert np.array_equal(sp, pos[idx])


# --------------------------------------------------------------------------
# collideSpheres (particles_kernel.cu:185-218): two-sphere closed forms
# --------------------------------------------------------------------------
def _P(**kw):
    return orc.make_params(2, 16, 0.5, **kw)


def test_pair_force_head_on_spring():
    P = _P()  # spring .5, damping .02, shear .1
    f = orc.pair_force([0, 0, 0], [0.75, 0, 0], [0, 0, 0], [0, 0, 0], 0.5, 0.5, 0.0, P)
    # -k (R - d) n with n from A to B: A is pushed away from B
    assert np.allclose(f, [-0.5 * 0.25, 0, 0], rtol=0, atol=1e-7)


def test_pair_force_diagonal_normal():
    P = _P()
    f = orc.pair_force([0, 0, 0], [0.3, 0.4, 0], [0, 0, 0], [0, 0, 0], 0.5, 0.5, 0.0, P)
    assert np.allclose(f, [-0.5 * 0.5 * 0.6, -0.5 * 0.5 * 0.8, 0], atol=1e-7)


def test_pair_force_normal_dashpot():
    P = _P()
    f = orc.pair_force([0, 0, 0], [0.75, 0, 0], [0, 0, 0], [0.5, 0, 0], 0.5, 0.5, 0.0, P)
    # relative velocity along n: no tangential part, dashpot c*u
    assert np.allclose(f, [-0.125 + 0.02 * 0.5, 0, 0], atol=1e-7)


def test_pair_force_tangential_shear():
    P = _P()
    f = orc.pair_force([0, 0, 0], [0.75, 0, 0], [0, 0, 0], [0, 0.5, 0], 0.5, 0.5, 0.0, P)
    # u = (0, .5, 0) is purely tangential: dashpot c*u plus shear s*u
    assert np.allclose(f, [-0.125, 0.02 * 0.5 + 0.1 * 0.5, 0], atol=1e-7)


def test_pair_force_velocity_of_A_enters_negatively():
    P = _P()
    f = orc.pair_force([0, 0, 0], [0.75, 0, 0], [0, 0, 0.5], [0, 0, 0], 0.5, 0.5, 0.0, P)
    assert np.allclose(f, [-0.125, 0, -(0.02 + 0.1) * 0.5], atol=1e-7)


def test_pair_force_attraction_uses_relpos():
    P = _P()
    f = orc.pair_force([0, 0, 0], [0.75, 0, 0], [0, 0, 0], [0, 0, 0], 0.5, 0.5, 0.3, P)
    assert np.allclose(f, [-0.125 + 0.3 * 0.75, 0, 0], atol=1e-7)


@pytest.mark.parametrize("d", [1.0, 1.0000001, 1.5])
def test_pair_force_zero_at_or_beyond_contact(d):
    P = _P()
    f = orc.pair_force([0, 0, 0], [d, 0, 0], [0, 0, 0], [1, 2, 3], 0.5, 0.5, 0.3, P)
    assert np.all(f == 0)


# --------------------------------------------------------------------------
# collideD (particles_kernel.cu:256-293): grid search vs all pairs
# --------------------------------------------------------------------------
def brute_force_newvel(pos, vel, P, r, collider=True):
    """All-pairs binary64 restatement: independent of cells and sorting."""
    p = pos[:, :3].astype(np.float64)
    v = vel[:, :3].astype(np.float64)
    k, c, s, a = (float(np.float32(x)) for x in (P.spring, P.damping, P.shear, P.attraction))

    def forces(pa, pb, va, vb, R, att):
        d = pb - pa
        dist = np.sqrt(np.sum(d * d, axis=-1))
        hit = dist < R
        with np.errstate(invalid="ignore", divide="ignore"):
            nrm = d / dist[..., None]
        u = vb - va
        ut = u - np.sum(u * nrm, axis=-1)[..., None] * nrm
        f = -k * (R - dist)[..., None] * nrm + c * u + s * ut + att * d
        return np.where(hit[..., None], f, 0.0)

    n = len(p)
    out = np.empty((n, 3))
    R = 2 * r
    for i in range(n):
        f = forces(p[i], p, v[i], v, R, a)
        f[i] = 0
        tot = f.sum(axis=0)
        if collider:
            cp = np.array(list(P.colliderPos), np.float64)
            tot += forces(p[i], cp, v[i], np.zeros(3), r + float(np.float32(P.colliderRadius)), 0.0)
        out[i] = v[i] + tot
    return out


def _grid_case(n, seed, grid=16, collider_pos=(0.2, -0.1, 0.3), attraction=0.0):
    r = float(np.float32(1.0 / grid))  # cell = 2r: every contact is within one cell
    P = orc.make_params(n, grid, r, colliderPos=collider_pos, colliderRadius=0.25,
                        attraction=attraction)
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-0.95, 0.95, size=(n, 3))
    pos[:, 3] = 1
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.uniform(-0.05, 0.05, size=(n, 3))
    h, idx = orc.calc_hash(pos, P)
    h, idx = orc.sort_pairs(h, idx)
    cs, ce, sp, sv = orc.reorder(h, idx, pos, vel, P.numCells)
    return P, r, pos, vel, (sp, sv, idx, cs, ce)


@pytest.mark.parametrize("seed,attr", [(0, 0.0), (1, 0.05)])
def test_collide_f64_equals_all_pairs(seed, attr):
    P, r, pos, vel, t = _grid_case(600, seed, attraction=attr)
    got = orc.collide_f64(*t, P)
    want = brute_force_newvel(pos, vel, P, r)
    assert np.allclose(got, want, rtol=0, atol=1e-13)


def test_collide_f32_close_to_f64():
    P, r, pos, vel, t = _grid_case(2000, 3)
    f32 = orc.collide(*t, P)
    f64 = orc.collide_f64(*t, P)
    force32 = f32[:, :3].astype(np.float64) - vel[:, :3]
    force64 = f64 - vel[:, :3]
    scale = np.max(np.abs(force64))
    assert scale > 1e-3  # the case has real contacts
    assert np.max(np.abs(force32 - force64)) <= 1e-5 * scale
    assert np.all(f32[:, 3] == 0)


def test_collide_conserves_momentum_without_external_forces():
    P, r, pos, vel, t = _grid_case(1500, 4, collider_pos=(50.0, 50.0, 50.0), attraction=0.02)
    f64 = orc.collide_f64(*t, P)
    dp = (f64 - vel[:, :3].astype(np.float64)).sum(axis=0)
    assert np.max(np.abs(f64 - vel[:, :3])) > 1e-3
    assert np.all(np.abs(dp) < 1e-12)


def test_collide_threads_bitwise_identical():
    P, r, pos, vel, t = _grid_case(3000, 5)
    orc.set_threads(1)
    a = orc.collide(*t, P)
    orc.set_threads(4)
    b = orc.collide(*t, P)
    orc.set_threads(1)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


# --------------------------------------------------------------------------
# integrate_functor (particles_kernel.cu:39-75)
# --------------------------------------------------------------------------
def test_integrate_ballistic_exact():
    P = orc.make_params(1, 64, 1.0 / 64.0, gravity=(0.0, 0.0, -0.25), globalDamping=1.0)
    pos = np.array([[0.25, 0.0, 0.0, 7.0]], np.float32)
    vel = np.array([[0.5, -0.25, 0.0, 3.0]], np.float32)
    p, v = orc.integrate(pos, vel, 0.5, P)
    assert v.tolist() == [[0.5, -0.25, -0.125, 3.0]]
    assert p.tolist() == [[0.5, -0.125, -0.0625, 7.0]]  # w lanes preserved


def test_integrate_global_damping_after_gravity():
    P = orc.make_params(1, 64, 1.0 / 64.0, gravity=(0.0, 0.0, -0.5), globalDamping=0.5)
    p, v = orc.integrate([[0, 0, 0, 1]], [[0, 0, 1.0, 0]], 1.0, P)
    # v = (1 - 0.5) * 0.5 = 0.25; p = 0 + 0.25 * 1
    assert v[0, 2] == 0.25 and p[0, 2] == 0.25


def test_integrate_walls_reflect_and_clamp():
    r = 1.0 / 64.0
    P = orc.make_params(6, 64, r, gravity=(0.0, 0.0, 0.0))
    pos = np.array([[0.99, 0, 0, 1], [-0.99, 0, 0, 1], [0, 0.99, 0, 1],
                    [0, -0.99, 0, 1], [0, 0, 0.99, 1], [0, 0, -0.99, 1]], np.float32)
    vel = np.zeros((6, 4), np.float32)
    for i in range(6):
        vel[i, i // 2] = 0.1 if i % 2 == 0 else -0.1
    p, v = orc.integrate(pos, vel, 0.5, P)
    for i in range(6):
        a = i // 2
        want = np.float32(1.0) - np.float32(r) if i % 2 == 0 else np.float32(-1.0) + np.float32(r)
        assert p[i, a] == want
        assert v[i, a] == np.float32(vel[i, a]) * np.float32(-0.5)
        others = [b for b in range(3) if b != a]
        assert np.all(p[i, others] == pos[i, others])


def test_integrate_corner_hits_every_wall():
    r = 1.0 / 64.0
    P = orc.make_params(1, 64, r, gravity=(0.0, 0.0, 0.0))
    p, v = orc.integrate([[0.99, -0.99, 0.99, 1]], [[0.1, -0.1, 0.1, 0]], 0.5, P)
    hi, lo = np.float32(1) - np.float32(r), np.float32(-1) + np.float32(r)
    assert p[0, :3].tolist() == [hi, lo, hi]
    assert np.allclose(v[0, :3], [-0.05, 0.05, -0.05], rtol=1e-7)


# --------------------------------------------------------------------------
# reset(CONFIG_GRID) and the observables
# --------------------------------------------------------------------------
def test_reset_grid_lattice_geometry():
    r = 1.0 / 64.0
    pos, vel = orc.reset(4096, 1, 1973, r, 0.1, 0.01)
    s = 16
    ijk = np.stack(np.meshgrid(np.arange(s), np.arange(s), np.arange(s), indexing="ij"), -1)
    ijk = ijk.transpose(2, 1, 0, 3).reshape(-1, 3)  # i = (z*s + y)*s + x
    centre = ijk * (2 * r) + r - 1
    assert np.all(np.abs(pos[:, :3] - centre) <= 0.1 * r + 1e-7)
    assert np.all(pos[:, 3] == 1) and np.all(vel[:, 3] == 0)
    assert np.all(np.abs(vel[:, :3]) <= 0.01)
    assert np.std(pos[:, :3] - centre) > 0.03 * r  # jitter present


def test_reset_partial_lattice_and_random():
    pos, _ = orc.reset(1000, 1, 7, 1 / 64, 0.0, 0.0)  # s = 10, exact cube
    assert len(np.unique(pos[:, 0])) == 10
    pos2, _ = orc.reset(1001, 1, 7, 1 / 64, 0.0, 0.0)  # s = 11, partial
    assert len(np.unique(pos2[:, 0])) == 11
    pr, vr = orc.reset(10000, 0, 7, 1 / 64, 0.0, 0.0)
    assert np.all(np.abs(pr[:, :3]) <= 1) and np.all(vr == 0)


def test_energy_momentum_compensated():
    rng = np.random.default_rng(9)
    vel = np.zeros((100_000, 4), np.float32)
    vel[:, :3] = rng.uniform(-1, 1, size=(100_000, 3))
    e = orc.energy_momentum(vel)
    v = vel[:, :3].astype(np.float64)
    assert abs(e[0] - 0.5 * np.sum(v * v)) <= 1e-9 * e[0]
    assert np.allclose(e[1:], v.sum(axis=0), rtol=0, atol=1e-9)


# --------------------------------------------------------------------------
# A full update (particleSystem.cpp:222-284)
# --------------------------------------------------------------------------
def test_update_step_invariants_and_thread_independence():
    r = 1.0 / 64.0
    P = orc.make_params(4096, 64, r)
    pos, vel = orc.reset(4096, 1, 1973, r, 0.1, 0.01)
    orc.set_threads(1)
    a = orc.State(pos, vel, P.numCells)
    a.run(0.5, 5, P)
    orc.set_threads(4)
    b = orc.State(pos, vel, P.numCells)
    b.run(0.5, 5, P)
    orc.set_threads(1)
    for name in ("pos4", "vel4", "hash", "index", "cellStart", "cellEnd", "sortedPos4"):
        assert np.array_equal(getattr(a, name).view(np.uint32), getattr(b, name).view(np.uint32))
    assert np.all(np.diff(a.hash.astype(np.int64)) >= 0)
    cell_table_invariants(a.hash, a.cellStart, a.cellEnd, 4096)
    assert np.array_equal(np.sort(a.index), np.arange(4096))

"""Ensemble-statistics helpers and the oracle's contact counts and FP64
observables, pinned by closed forms on the CPU."""
from __future__ import annotations

import math

import numpy as np

from fullc0ntact_amd import statistics as S
from oracle import oracle as orc


def _lattice(s, spacing, r, origin=-1.0):
    g = np.stack(np.meshgrid(*[np.arange(s)] * 3, indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
    pos = np.ones((len(g), 4), np.float32)
    pos[:, :3] = origin + r + g * spacing
    return pos


def test_packing_fraction_simple_cubic_is_pi_over_6():
    r = 1.0 / 16
    pos = _lattice(16, 2 * r, r)  # fills [-1, 1]^3 exactly, 16 layers
    pos[:, 3] = r
    phi = S.packing_fraction(pos)
    assert abs(phi - math.pi / 6) < 1e-12


def test_packing_fraction_uses_quantile_height():
    r = 1.0 / 16
    pos = _lattice(16, 2 * r, r)
    pos[:, 3] = r
    # lift 2% of the spheres high above the bed: the 95% bed is unchanged
    rng = np.random.default_rng(0)
    idx = rng.choice(len(pos), len(pos) // 50, replace=False)
    pos[idx, 2] += 5.0
    phi = S.packing_fraction(pos)
    assert 0.45 < phi < math.pi / 6 + 1e-12


def test_polydisperse_radii_distribution():
    med = 1.0 / 512
    r = S.polydisperse_radii(200000, med, 0.25, 3.0, seed=7)
    assert r.dtype == np.float32
    assert r.max() / r.min() <= 3.0 * (1 + 1e-6)
    assert abs(np.median(r) / med - 1.0) < 0.01
    # the unclipped part is a normal in log r truncated at +-ln(sqrt 3)
    from scipy.stats import truncnorm
    a = np.log(np.sqrt(3.0)) / 0.25
    inner = np.log(r[(r > r.min()) & (r < r.max())].astype(np.float64))
    assert abs(np.std(inner) - 0.25 * truncnorm(-a, a).std()) < 0.003


def _oracle_tables(pos, P):
    h, i = orc.calc_hash(pos, P)
    h, i = orc.sort_pairs(h, i)
    cs, ce, sp, sv = orc.reorder(h, i, pos, np.zeros_like(pos), P.numCells)
    return sp, cs, ce


def test_contact_counts_closed_form_lattice():
    """Spacing 1.9r: face neighbours touch (1.9r < 2r), diagonals do not
    (2.69r): 6 contacts inside, 5 on faces, 4 on edges, 3 at corners."""
    s, r = 10, 1.0 / 64
    P = orc.make_params(s ** 3, 64, r, colliderPos=(5.0, 5.0, 5.0))
    pos = _lattice(s, np.float32(1.9 * r), r)
    sp, cs, ce = _oracle_tables(pos, P)
    hist = orc.contact_histogram(sp, cs, ce, P, 8)
    expect = np.zeros(8, np.uint32)
    expect[3] = 8
    expect[4] = 12 * (s - 2)
    expect[5] = 6 * (s - 2) ** 2
    expect[6] = (s - 2) ** 3
    assert np.array_equal(hist, expect)
    # touching at exactly 2r is not a contact (strict test)
    pos2 = _lattice(s, np.float32(2 * r), r)
    sp, cs, ce = _oracle_tables(pos2, P)
    assert orc.contact_histogram(sp, cs, ce, P, 8)[0] == s ** 3


def test_contact_counts_polydisperse_pairs():
    """Two spheres at distance d touch iff d < rA + rB."""
    P = orc.make_params(2, 64, 1.0 / 32)
    for ra, rb, d, touch in [(0.02, 0.01, 0.0299, True), (0.02, 0.01, 0.0301, False),
                             (0.005, 0.03, 0.034, True)]:
        pos = np.array([[0.0, 0.0, 0.0, ra], [d, 0.0, 0.0, rb]], np.float32)
        sp, cs, ce = _oracle_tables(pos, P)
        c = orc.contact_counts(sp, cs, ce, P, poly=True)
        assert list(c) == ([1, 1] if touch else [0, 0])


def test_energy_momentum_oracle_exact_sums():
    rng = np.random.default_rng(3)
    vel = np.zeros((100003, 4), np.float32)
    vel[:, :3] = rng.normal(0, 0.01, size=(len(vel), 3))
    out = orc.energy_momentum(vel)
    v = vel[:, :3].astype(np.float64)
    ke = math.fsum(0.5 * (v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2]))
    assert abs(out[0] - ke) <= 1e-15 * ke
    for a in range(3):
        assert abs(out[1 + a] - math.fsum(v[:, a])) <= 1e-15 * np.abs(v[:, a]).sum()


def test_polydisperse_equal_radii_reduce_to_monodisperse():
    n, r = 4096, 1.0 / 64
    P = orc.make_params(n, 64, r)
    pos, vel = orc.reset(n, 1, 1973, r, 0.1, 0.01)
    pos[:, 3] = np.float32(r)
    a = orc.State(pos, vel, P.numCells)
    b = orc.State(pos, vel, P.numCells)
    for _ in range(10):
        a.update(0.5, P)
        orc.update_ex(b, 0.5, P, None, poly=True)
    assert np.array_equal(a.pos4.view(np.uint32), b.pos4.view(np.uint32))
    assert np.array_equal(a.vel4.view(np.uint32), b.vel4.view(np.uint32))

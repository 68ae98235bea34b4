"""Pins the oracle's signed-distance MLP (fc_oracle_sdf.c) on the CPU:
closed-form forwards of hand-set networks, central finite differences of
the loss for the analytic backward, and the BF16 emulation's closeness."""
from __future__ import annotations

import numpy as np

from oracle import oracle as orc

NP = orc.SDF_NPARAMS
OW1, OB1, OW2, OB2, OW3, OB3, OW4, OB4 = 0, 2048, 2112, 6208, 6272, 10368, 10432, 10496
RNG = np.random.default_rng(31)


def _net(scale=0.3):
    p = np.zeros(NP, np.float32)
    p[:] = RNG.uniform(-scale, scale, NP)
    freq = RNG.normal(0, 1.0, (16, 3)).astype(np.float32)
    frame = np.array([0.1, -0.2, 0.05, 0.7], np.float32)
    return p, freq, frame


def _silu(z):
    return z / (1 + np.exp(-z))


def test_forward_constant_and_single_path():
    p, freq, frame = _net()
    q = RNG.uniform(-0.5, 0.5, (50, 3)).astype(np.float32)
    c = np.zeros(NP, np.float32)
    c[OB4] = 0.25
    assert np.allclose(orc.sdf_forward(c, freq, frame, q), 0.25 * frame[3], rtol=0, atol=1e-15)
    s = np.zeros(NP, np.float32)
    s[OW1 + 0 * 64 + 0] = 1.0
    s[OW2 + 0 * 64 + 0] = 1.0
    s[OW3 + 0 * 64 + 0] = 1.0
    s[OW4 + 0] = 1.0
    x = (q.astype(np.float64) - frame[:3]) / frame[3]
    t = x @ freq[0].astype(np.float64)
    expect = _silu(_silu(_silu(np.sin(2 * np.pi * t)))) * frame[3]
    assert np.allclose(orc.sdf_forward(s, freq, frame, q), expect, rtol=1e-12, atol=1e-14)
    # the cosine features are inputs 16..31
    s2 = np.zeros(NP, np.float32)
    s2[OW1 + 16 * 64 + 5] = 1.0
    s2[OW2 + 5 * 64 + 7] = 1.0
    s2[OW3 + 7 * 64 + 9] = 1.0
    s2[OW4 + 9] = 1.0
    expect2 = _silu(_silu(_silu(np.cos(2 * np.pi * t)))) * frame[3]
    assert np.allclose(orc.sdf_forward(s2, freq, frame, q), expect2, rtol=1e-12, atol=1e-14)


def test_loss_gradient_by_finite_differences():
    p, freq, frame = _net()
    q = RNG.uniform(-0.6, 0.6, (40, 3)).astype(np.float32)
    lab = RNG.uniform(-0.3, 0.3, 40).astype(np.float32)
    g, L = orc.sdf_loss_grad(p, freq, frame, q, lab)
    y = orc.sdf_forward(p, freq, frame, q) / frame[3]
    assert abs(L - np.mean((y - lab.astype(np.float64) / float(frame[3])) ** 2)) < 1e-12
    idx = np.concatenate([RNG.choice(np.arange(s, e), 6, replace=False) for s, e in
                          [(OW1, OB1), (OB1, OW2), (OW2, OB2), (OB2, OW3), (OW3, OB3),
                           (OB3, OW4), (OW4, OB4)]] + [np.array([OB4])])
    h = np.float32(2.0 ** -12)
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        hp = float(pp[i]) - float(p[i])
        hm = float(p[i]) - float(pm[i])
        _, Lp = orc.sdf_loss_grad(pp, freq, frame, q, lab)
        _, Lm = orc.sdf_loss_grad(pm, freq, frame, q, lab)
        fd = (Lp - Lm) / (hp + hm)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(g[i]), (i, fd, g[i])


def test_bf16_emulation_is_close():
    p, freq, frame = _net(0.2)
    q = RNG.uniform(-0.6, 0.6, (200, 3)).astype(np.float32)
    a = orc.sdf_forward(p, freq, frame, q)
    b = orc.sdf_forward(p, freq, frame, q, bf16=True)
    assert np.max(np.abs(a - b)) <= 0.03 * np.max(np.abs(a)) + 1e-6
    assert not np.array_equal(a, b)

"""The row-range collide's candidate filter on 16-bit fixed-point coordinates
must keep every pair that the reference counts as touching.

The device arithmetic is emulated bit for bit on the host: the coordinate
u = rint((p - lo) * scale) in float32 (separate rounding of the difference
and the product, as compiled without contraction), clamped to [0, 65535];
the x and y differences in wrapping 16-bit lanes, the z difference exact;
the squares summed modulo 2^32.  The
reference's contact test is float32 `sqrtf(dot(rel, rel)) < rA + rB`.  Pairs
are placed at the contact distance to the last ulp, in random directions
(axis-aligned and diagonal ones included), all over the frame and beyond
it, where the coordinates clamp.  The bound (2r * scale + 2.5)^2 must hold
for all of them, and must be tight enough that the filter still drops
pairs a few per cent beyond contact.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

BOX = (-1.0, 65535.0 / 2.0)      # lo, scale of the box walls' frame
MAP = (-1.0, 65535.0 / 11.0)     # lo, scale with a distance-map boundary


def threshold(R, scale):
    b = float(np.float32(R)) * float(np.float32(scale)) + 2.5
    return 0xFFFFFFFF if not b < 32000.0 else math.ceil(b * b)


def quantize(p, lo, scale):
    u = np.rint((p - np.float32(lo)).astype(np.float32) * np.float32(scale)).astype(np.float32)
    return np.fmin(np.fmax(u, np.float32(0.0)), np.float32(65535.0)).astype(np.int64)


def quant_d2(ua, ub):
    dxy = (ua[:, :2] - ub[:, :2] + 32768) % 65536 - 32768  # int16 lanes
    dz = ua[:, 2] - ub[:, 2]                                # exact
    return ((dxy * dxy).sum(axis=1) + dz * dz) % (1 << 32)


def ref_touch(pa, pb, cd):
    rel = (pb - pa).astype(np.float32)
    d2 = (rel[:, 0] * rel[:, 0] + rel[:, 1] * rel[:, 1]).astype(np.float32)
    d2 = (d2 + rel[:, 2] * rel[:, 2]).astype(np.float32)
    return np.sqrt(d2).astype(np.float32) < cd


def pairs(rng, n, r, extent, spread_ulps=6):
    """Pairs at float32 separations within a few ulp of their contact
    distance rA + rB (radii in [r/2, r])."""
    ra = (rng.uniform(0.5, 1.0, n) * r).astype(np.float32)
    rb = (rng.uniform(0.5, 1.0, n) * r).astype(np.float32)
    cd = (ra + rb).astype(np.float32)
    u = rng.normal(size=(n, 3))
    k = n // 8  # axis-aligned and diagonal directions too
    u[:k] = np.eye(3)[rng.integers(0, 3, k)] * rng.choice([-1.0, 1.0], (k, 1))
    u[k:2 * k] = rng.choice([-1.0, 1.0], (k, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    d = cd.astype(np.float64) * (1.0 + rng.integers(-spread_ulps, spread_ulps + 1, n) * 2.0 ** -24)
    pa = rng.uniform(extent[0], extent[1], (n, 3)).astype(np.float32)
    pb = (pa + (u * d[:, None]).astype(np.float32)).astype(np.float32)
    return pa, pb, cd


@pytest.mark.parametrize("frame,r,extent", [
    (BOX, 1.0 / 128, (-1.0, 1.0)),
    (BOX, 1.0 / 64, (-1.2, 1.2)),       # beyond the walls: clamped coordinates
    (BOX, 1.0 / 512, (-1.0, 1.0)),
    (MAP, 1.0 / 512 * math.sqrt(3.0), (-1.0, 10.0)),
    (MAP, 1.0 / 128, (-3.0, 12.0)),
])
def test_filter_keeps_every_touching_pair(frame, r, extent):
    lo, scale = frame
    rng = np.random.default_rng(7)
    T = threshold(2 * r, scale)
    kept_total = touching_total = 0
    for _ in range(8):
        pa, pb, cd = pairs(rng, 250_000, r, extent)
        touch = ref_touch(pa, pb, cd)
        d2 = quant_d2(quantize(pa, lo, scale), quantize(pb, lo, scale))
        kept = d2 < T
        assert np.all(kept[touch]), "a touching pair was dropped"
        kept_total += int(kept.sum())
        touching_total += int(touch.sum())
    assert touching_total > 100_000


def test_filter_drops_pairs_beyond_its_margin():
    # Inside the frame the bound is tight: pairs 2 % beyond the largest
    # contact distance (r = 1/128, the 1M-sphere workload) are dropped.
    lo, scale = BOX
    r = 1.0 / 128
    rng = np.random.default_rng(8)
    T = threshold(2 * r, scale)
    n = 200_000
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    pa = rng.uniform(-0.98, 0.98, (n, 3)).astype(np.float32)
    pb = (pa + (u * (2 * r * 1.02)).astype(np.float32)).astype(np.float32)
    d2 = quant_d2(quantize(pa, lo, scale), quantize(pb, lo, scale))
    assert not np.any(d2 < T)


def test_nan_coordinates_quantize_to_zero():
    # fmaxf(NaN, 0) is 0 on the device: a NaN sphere gets coordinate 0 and
    # its pairs reach the exact test, which gives them +0.
    lo, scale = BOX
    u = quantize(np.array([[np.nan, 0.0, 1.0]], np.float32), lo, scale)
    assert u.tolist() == [[0, 32768, 65535]]

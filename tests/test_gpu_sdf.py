"""The signed-distance surrogate on the GPU against the oracle.

 - MFMA lane layout: the probe's 16x32 by 32x16 BF16 product is exact;
 - FP32 forward and loss gradient vs the binary64 oracle (the gradient is
   the GPU's own backward kernel; 1e-4 of the largest component);
 - BF16 MFMA forward vs the oracle's BF16 emulation and vs FP32;
 - training: deterministic, the loss falls, and the trained network's BF16
   distances track the exact mesh distance.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from fullc0ntact_amd import meshes
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _threads():
    orc.set_threads(max(1, min(16, os.cpu_count() or 1)))
    yield
    orc.set_threads(1)


def _bf16(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def test_mfma_probe_layout(gpu_ctx):
    from fullc0ntact_amd import _lib
    rng = np.random.default_rng(4)
    A = rng.integers(-4, 5, (16, 32)).astype(np.float32)
    B = rng.integers(-4, 5, (32, 16)).astype(np.float32)
    da = _lib.DeviceArray.from_host(gpu_ctx, _bf16(A))
    db = _lib.DeviceArray.from_host(gpu_ctx, _bf16(B))
    dd = _lib.DeviceArray(gpu_ctx, 16 * 16 * 4)
    _lib.check(_lib.lib.fc_mfma_probe(gpu_ctx.handle, da.ptr, db.ptr, dd.ptr))
    D = dd.download(np.float32, 256).reshape(16, 16)
    assert np.array_equal(D, A @ B)


def _surrogate(gpu_ctx, nb=2, seed=7):
    from fullc0ntact_amd.sdf import SdfSurrogate
    centers = np.array([[0.1, -0.05, 0.0], [0.0, 0.2, -0.1], [0.3, 0.3, 0.3]][:nb], np.float32)
    half = np.array([0.6, 0.8, 0.5][:nb], np.float32)
    return SdfSurrogate(gpu_ctx, centers, half, freq_sigma=1.0, seed=seed)


def _perturb(s, scale=0.05, seed=3):
    p, f, fr = s.params()
    rng = np.random.default_rng(seed)
    p[:, :s.nparams] += rng.uniform(-scale, scale, (s.num_bodies, s.nparams)).astype(np.float32)
    s.set_params(p)
    return s.params()


def test_forward_fp32_and_bf16_match_oracle(gpu_ctx):
    from fullc0ntact_amd.sdf import BF16, FP32
    s = _surrogate(gpu_ctx, 3)
    p, f, fr = _perturb(s)
    rng = np.random.default_rng(9)
    n = 5003  # ragged: the last MFMA tile is partial
    pts = rng.uniform(-0.8, 0.8, (n, 3)).astype(np.float32)
    bodies = rng.integers(0, 3, n).astype(np.uint32)
    bodies[:1000] = 1  # long single-body runs and mixed tiles
    g32 = s.eval(pts, bodies, FP32)
    g16 = s.eval(pts, bodies, BF16)
    for b in range(3):
        m = bodies == b
        ref = orc.sdf_forward(p[b], f[b], fr[b], pts[m])
        ref16 = orc.sdf_forward(p[b], f[b], fr[b], pts[m], bf16=True)
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(g32[m] - ref)) <= 1e-4 * scale, b
        assert np.max(np.abs(g16[m] - ref16)) <= 2e-3 * scale, b
        assert np.max(np.abs(g16[m] - ref)) <= 3e-2 * scale, b
    # body ids out of range give NaN, others are unaffected
    bad = bodies.copy()
    bad[::7] = 99
    g = s.eval(pts, bad, BF16)
    assert np.all(np.isnan(g[::7]))
    ok = np.ones(n, bool)
    ok[::7] = False
    assert np.array_equal(g[ok], g16[ok])
    s.close()


def test_loss_gradient_matches_oracle(gpu_ctx):
    from fullc0ntact_amd import _lib
    s = _surrogate(gpu_ctx, 2)
    p, f, fr = _perturb(s, 0.03)
    rng = np.random.default_rng(12)
    m = 700  # not a multiple of the 64-sample tile
    pts = rng.uniform(-0.7, 0.7, (2, m, 3)).astype(np.float32)
    lab = rng.uniform(-0.3, 0.3, (2, m)).astype(np.float32)
    dp = _lib.DeviceArray.from_host(gpu_ctx, pts)
    dl = _lib.DeviceArray.from_host(gpu_ctx, lab)
    g, loss = s.loss_grad(dp, dl, m)
    for b in range(2):
        ref, L = orc.sdf_loss_grad(p[b], f[b], fr[b], pts[b], lab[b])
        gb = g[b, :s.nparams].astype(np.float64)
        assert abs(loss[b] - L) <= 1e-5 * L
        assert np.max(np.abs(gb - ref)) <= 1e-4 * np.max(np.abs(ref)), b
    g2, _ = s.loss_grad(dp, dl, m)
    assert np.array_equal(g, g2)  # deterministic
    s.close()


def _sphere_set(gpu_ctx, n, center, R, half, seed):
    from fullc0ntact_amd import _lib
    from fullc0ntact_amd.distance_map import mesh_signed_distance
    from fullc0ntact_amd.sdf import training_samples
    v, t = meshes.uv_sphere(R, 48, 24, center=center)
    pts = training_samples(v, t, n, center, half, 0.5, 0.05, seed=seed)
    sd, _ = mesh_signed_distance(gpu_ctx, v, t, pts)
    return pts.astype(np.float32), sd.astype(np.float32)


def test_training_is_deterministic_and_learns(gpu_ctx):
    from fullc0ntact_amd import _lib
    from fullc0ntact_amd.sdf import BF16, SdfSurrogate
    n = 1 << 14
    centers = np.array([[0.0, 0.0, 0.0], [0.2, -0.1, 0.1]], np.float32)
    R = [0.3, 0.2]
    half = np.array([0.5, 0.35], np.float32)
    sets = [_sphere_set(gpu_ctx, n, centers[b], R[b], half[b], 20 + b) for b in range(2)]
    pts = np.stack([s[0] for s in sets])
    lab = np.stack([s[1] for s in sets])
    dp = _lib.DeviceArray.from_host(gpu_ctx, pts)
    dl = _lib.DeviceArray.from_host(gpu_ctx, lab)
    runs = []
    for _ in range(2):
        s = SdfSurrogate(gpu_ctx, centers, half, 1.0, seed=5)
        hist = s.train(dp, dl, n, 512, 60, 3e-3, 0.98)
        runs.append((s.params()[0], hist))
        last = s
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    hist = runs[0][1]
    print("loss first/last epoch", hist[0], hist[-1], flush=True)
    assert np.all(hist[-1] < 0.02 * hist[0])
    # held-out exact distances inside the normalisation cube (queries
    # outside a body's box never reach its network) are reproduced
    rng = np.random.default_rng(99)
    for b in range(2):
        q = centers[b] + rng.uniform(-1.0, 1.0, (4000, 3)) * half[b]
        q = q.astype(np.float32)
        exact = np.linalg.norm(q - centers[b], axis=1) - R[b]
        pred = last.eval(q, np.full(len(q), b, np.uint32), BF16)
        err = np.abs(pred - exact)
        print(f"body {b}: mean |err| {err.mean():.2e}, max {err.max():.2e}", flush=True)
        assert np.mean(err) < 0.02 * half[b], b
    last.close()

"""fc_sort_particles_noalloc: the (hash, index) sort without any device
allocation (scratch in one module array per device), for callers that cannot
call hipMalloc.  Checked through the C ABI against fc_sort_particles (and the
oracle's (hash, id) order) on two calls in a row, the second with a smaller
input, so the second call runs on the scratch the first one left behind."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _keys(n, grid, seed):
    rng = np.random.default_rng(seed)
    # cells of a 256^3 grid with many duplicates, ids in order
    h = rng.integers(0, grid ** 3 // 64, n, dtype=np.uint64).astype(np.uint32) * 64
    h += rng.integers(0, 3, n).astype(np.uint32)
    return h, np.arange(n, dtype=np.uint32)


def _sort(ctx, fn, h, i):
    from fullc0ntact_amd import _lib
    n = len(h)
    dh = _lib.DeviceArray.from_host(ctx, h)
    di = _lib.DeviceArray.from_host(ctx, i)
    _lib.check(fn(ctx.handle, dh.ptr, di.ptr, n))
    ctx.sync()
    return dh.download(np.uint32, n), di.download(np.uint32, n)


def test_noalloc_sort_equals_sort_on_consecutive_calls(gpu_ctx):
    from fullc0ntact_amd import _lib, make_params
    L = _lib.lib
    assert L.fc_sort_particles_noalloc_max() == 1 << 20
    gpu_ctx.set_parameters(make_params(0, 256, 1.0 / 128))
    for n, seed in ((1 << 20, 3), (100_003, 4), (4097, 5), (1, 6)):
        print(f"[noalloc] n={n}", flush=True)
        h, i = _keys(n, 256, seed)
        ah, ai = _sort(gpu_ctx, L.fc_sort_particles_noalloc, h, i)
        bh, bi = _sort(gpu_ctx, L.fc_sort_particles, h, i)
        assert np.array_equal(ah, bh) and np.array_equal(ai, bi)
        oh, oi = orc.sort_pairs(h, i)
        assert np.array_equal(ah, oh) and np.array_equal(ai, oi)
    # beyond the stated maximum: a clear error, nothing launched
    with pytest.raises(_lib.FcError, match="at most 1048576"):
        h, i = _keys((1 << 20) + 1, 256, 7)
        _sort(gpu_ctx, L.fc_sort_particles_noalloc, h, i)

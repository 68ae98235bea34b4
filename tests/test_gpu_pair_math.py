"""The contact loop's fast square root and reciprocal are the correctly
rounded operations on every argument of their ranges.

The row-range collide takes sqrt(d2) through v_sqrt_f32 plus a neighbour
selection, and 1/dist through v_rcp_f32 plus one Newton step, whenever every
lane of the wave has d2 in [2^-96, 2^96] and dist's significand is not all
ones (else the general sequences).  fc_selftest_pair_math compares both
against the device's correctly rounded sqrtf and 1.0f / x on all 1.6e9 and
8.1e8 floats of the two ranges.  The bit-exact parity tests against the
oracle then cover the switch between the paths.
"""
from __future__ import annotations

import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def test_fast_sqrt_and_reciprocal_are_correctly_rounded(gpu_ctx):
    from fullc0ntact_amd import _lib
    out = (C.c_uint64 * 5)()
    _lib.check(_lib.lib.fc_selftest_pair_math(gpu_ctx.handle, out))
    print(f"sqrt mismatches {out[0]}, reciprocal mismatches {out[1]}, "
          f"all-ones significand reciprocal mismatches {out[2]} (taken by the general "
          f"division), first bad sqrt {out[3]:#x}, first bad reciprocal {out[4]:#x}", flush=True)
    assert out[0] == 0 and out[3] == 0xFFFFFFFF
    assert out[1] == 0 and out[4] == 0xFFFFFFFF

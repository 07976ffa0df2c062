"""The allocation-free shadow bisection (mk_shadow_bisect_noalloc) against
mk_shadow_bisect through the C ABI: the same radii bit for bit, on two calls
in a row that share the library's fixed staging array (the second smaller,
so stale contents from the first lie beyond its angles), and a clear error
beyond the size limit."""
import ctypes as C

import numpy as np
import pytest

from mahakala_amd import _lib

pytestmark = pytest.mark.gpu
MAX_ANGLES = 65536


def _bisect(fn, a, inc, angles, max_steps=2000, error_allowed=1e-3, max_it=40, distance=1000.0):
    ang = np.ascontiguousarray(angles, dtype=np.float64)
    out = np.full(ang.size, np.nan)
    _lib.check(fn(a, inc, ang.ctypes.data, ang.size, max_steps, error_allowed, max_it, distance, out.ctypes.data))
    return out


def test_noalloc_bisection_equals_the_allocating_entry_point():
    L = _lib.lib()
    cases = [(0.9, 60.0, np.arange(4096) / 4096 * 2 * np.pi),        # 4 rounds of 4 halvings
             (-0.5, 20.0, np.linspace(0.1, 6.0, 64), 1500, 1e-2)]    # smaller, 7 halvings in one round
    for case in cases:
        a, inc, ang = case[:3]
        extra = case[3:]
        ref = _bisect(L.mk_shadow_bisect, a, inc, ang, *extra)
        got = _bisect(L.mk_shadow_bisect_noalloc, a, inc, ang, *extra)
        print(f"[noalloc] a={a} i={inc} {ang.size} angles: max |diff| {np.max(np.abs(got - ref)):.1e}", flush=True)
        assert np.all(np.isfinite(ref)) and np.all((ref > 2) & (ref < 8))
        np.testing.assert_array_equal(got, ref)
    # the same call again: the staging array keeps the previous call's
    # contents, which must not leak into this one
    again = _bisect(L.mk_shadow_bisect_noalloc, 0.9, 60.0, cases[0][2][::64])
    np.testing.assert_array_equal(again, _bisect(L.mk_shadow_bisect, 0.9, 60.0, cases[0][2][::64]))


def test_noalloc_bisection_limits():
    L = _lib.lib()
    big = np.zeros(MAX_ANGLES + 1)
    out = np.empty_like(big)
    rc = L.mk_shadow_bisect_noalloc(0.5, 30.0, big.ctypes.data, big.size, 2000, 1e-3, 40, 1000.0, out.ctypes.data)
    assert rc == -1 and b"exceed the limit of 65536" in L.mk_last_error()
    assert L.mk_shadow_bisect_noalloc(0.5, 30.0, big.ctypes.data, 0, 2000, 1e-3, 40, 1000.0, out.ctypes.data) == 0

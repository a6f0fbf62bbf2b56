"""The allocation-free variant of the math hook (lemhip_math_eval_noalloc)
against lemhip_math_eval, through the C-ABI: bit-identical on two calls in a
row, the second with a smaller input (so nothing stale from the first can
leak into it), at the size limit, and refused above it."""
import ctypes as C

import numpy as np
import pytest

from parallem_amd import lemhip

pytestmark = pytest.mark.gpu

MAX = 65536


def _noalloc(fn, x, y):
    L = lemhip.lib()
    L.lemhip_math_eval_noalloc.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    out = np.full_like(x, np.nan)
    rc = L.lemhip_math_eval_noalloc(fn, x.size, x.ctypes.data, y.ctypes.data, out.ctypes.data)
    return rc, out


def test_noalloc_equals_allocating_hook_on_two_calls():
    rng = np.random.default_rng(17)
    for fn in range(6):
        lo, hi = ((1e-6, 1e4) if fn in (1, 2, 3) else (-40.0, 40.0))
        for n in (50000, 1000):
            x = rng.uniform(lo, hi, n)
            y = rng.uniform(-3, 3, n)
            print(f"[noalloc] fn {fn} n {n}", flush=True)
            rc, got = _noalloc(fn, x, y)
            assert rc == 0
            ref = lemhip.math_eval(fn, x, y)
            assert np.array_equal(got, ref, equal_nan=True), f"fn {fn} n {n}"


def test_noalloc_size_limit():
    x = np.linspace(-1, 1, MAX)
    rc, got = _noalloc(0, x, x)
    assert rc == 0 and np.array_equal(got, lemhip.math_eval(0, x, x))
    big = np.zeros(MAX + 1)
    rc, got = _noalloc(0, big, big)
    assert rc == -6 and np.isnan(got).all()
    rc, _ = _noalloc(0, np.zeros(0), np.zeros(0))
    assert rc == -1

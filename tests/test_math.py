"""The portable FP64 routines (include/lem_math.h) against the platform libm.

They replace the device math library on the reference's call sites; the
bound each test states is the accuracy they must keep.  Bit-identity of the
host and gfx950 builds is checked in the GPU parity suite.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def ulp_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    sp = np.spacing(np.abs(ref))
    return np.abs(got - ref) / np.where(sp > 0, sp, np.finfo(float).tiny)


def _ref(fn, xs, ys=None):
    if ys is None:
        return np.array([fn(x) for x in xs])
    return np.array([fn(x, y) for x, y in zip(xs, ys)])


@pytest.fixture(scope="module")
def rng():
    return np.random.default_rng(1234)


def test_exp(rng):
    x = np.concatenate([rng.uniform(-700, 700, 20000), rng.uniform(-1, 1, 5000), [0.0, 1e-300, -1e-300]])
    assert ulp_err(O.math_eval(0, x), _ref(math.exp, x)).max() <= 2.0


def test_exp_special():
    got = O.math_eval(0, np.array([np.nan, np.inf, -np.inf, 710.0, -746.0, 0.0]))
    assert np.isnan(got[0]) and got[1] == np.inf and got[2] == 0.0 and got[3] == np.inf and got[4] == 0.0
    assert got[5] == 1.0


def test_log_and_log10(rng):
    x = np.concatenate([np.exp(rng.uniform(-700, 700, 20000)), rng.uniform(0.5, 2.0, 5000), [1.0, 2.0, 10.0, 5e-324]])
    assert ulp_err(O.math_eval(1, x), _ref(math.log, x)).max() <= 2.0
    assert ulp_err(O.math_eval(2, x), _ref(math.log10, x)).max() <= 3.0
    got = O.math_eval(1, np.array([0.0, -1.0, np.inf, np.nan]))
    assert got[0] == -np.inf and np.isnan(got[1]) and got[2] == np.inf and np.isnan(got[3])


def test_pow_on_the_reference_call_sites(rng):
    # bases and exponents as they occur: fines^1.14, slope^1.67, (110-stone)^0.877,
    # f^1.3, 10^k for the Dreybrodt constants, x^(1/3)
    xs = np.concatenate([rng.uniform(0.1, 100, 3000), rng.uniform(1e-4, 5, 3000), rng.uniform(10, 110, 3000),
                         rng.uniform(0, 1e4, 3000), np.full(3000, 10.0), rng.uniform(1e-12, 1e-3, 3000)])
    ys = np.concatenate([np.full(3000, 1.14), np.full(3000, 1.67), np.full(3000, 0.877), np.full(3000, 1.3),
                         rng.uniform(-12, 8, 3000), np.full(3000, 1 / 3)])
    err = ulp_err(O.math_eval(3, xs, ys), _ref(math.pow, xs, ys))
    assert err.max() <= 3.0, err.max()


def test_pow_special():
    x = np.array([0.0, 0.0, 5.0, -2.0, -3.0, np.nan, 2.0])
    y = np.array([1.3, 0.0, 0.0, 0.5, 2.0, 1.0, np.nan])
    got = O.math_eval(3, x, y)
    assert got[0] == 0.0 and got[1] == 1.0 and got[2] == 1.0 and np.isnan(got[3])
    assert got[4] == 9.0 and np.isnan(got[5]) and np.isnan(got[6])


def test_tanh_and_expm1(rng):
    x = np.concatenate([rng.uniform(-25, 25, 20000), rng.uniform(-1e-6, 1e-6, 2000), [0.0, 30.0, -30.0]])
    assert ulp_err(O.math_eval(4, x), _ref(math.tanh, x)).max() <= 3.0
    assert ulp_err(O.math_eval(5, x), _ref(math.expm1, x)).max() <= 3.0
    assert np.isnan(O.math_eval(4, np.array([np.nan])))[0]

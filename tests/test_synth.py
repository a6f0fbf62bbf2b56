"""Synthetic workload generator: determinism and the rules of each workload DEM."""
import math

import numpy as np

from parallem_amd import synth


def test_deterministic():
    a = synth.fbm_512(side=64)
    b = synth.fbm_512(side=64)
    assert np.array_equal(a.z, b.z) and np.array_equal(a.rain, b.rain)


def test_fbm_512_rules():
    d = synth.fbm_512()
    z = d.z
    assert z.shape == (512, 512)
    f = math.ceil(0.05 * 512)
    assert np.all(z[:f] == -9999) and np.all(z[:, -f:] == -9999)
    inner = z[f:-f, f:-f]
    assert np.all(inner != -9999)
    assert inner.min() > -20 and inner.max() < 800 + 0.15 * 512 + 1
    assert np.all(d.rain == 1000.0)
    # plateaus: exact equal-height 8-connected groups exist
    eq = (inner[:, 1:] == inner[:, :-1]).sum()
    assert eq > 50


def test_fbm_4096_rain_lognormal_mean():
    d = synth.fbm_4096(side=256)
    m = d.z != -9999
    assert abs(d.rain[m].mean() - 1000.0) < 1e-6
    assert d.rain[m].std() > 100


def test_outline_batch():
    ds = synth.catchment_batch(3, seed=5, lo=64, hi=128)
    for d in ds:
        h, w = d.z.shape
        assert 64 <= h <= 128 and h == w
        assert np.all(d.z[0] == -9999) and np.all(d.z[:, 0] == -9999)
        assert (d.z != -9999).mean() > 0.2

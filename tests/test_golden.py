"""Golden fixtures (tests/golden/, written by tests/golden/make_fixtures.py
from the oracle once it was pinned by tests/test_oracle*.py): the oracle must
still produce every stored field and scalar bit for bit, and the digests of
the 512^2 runs.  A change to the oracle that moves any bit fails here."""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from parallem_amd import synth

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLD, "manifest.json")) as fh:
    MANIFEST = json.load(fh)
FIELDS_F = ["dem", "soilT", "stone", "fines", "fa", "depo", "soilM", "TotB"]
FIELDS_I = ["fd", "SFD", "mask"]


def load(name):
    with np.load(os.path.join(GOLD, f"{name}.npz")) as z:
        return {k: z[k] for k in z.files}


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", sorted(MANIFEST["small"]))
def test_oracle_reproduces_small_fixture(name):
    m = MANIFEST["small"][name]
    g = load(name)
    H, W = g["in_dem"].shape
    orc = O.Oracle(g["in_dem"], g["in_rain"], O.default_params(W, H, m["flowtype"], runiniter=m["runiniter"]),
                   soilT_mode=m["soilT_mode"])
    for k in range(m["steps"]):
        orc.step(-1000 + k)
    for f in FIELDS_F + FIELDS_I:
        assert np.array_equal(orc[f], g[f], equal_nan=True), f"{name}: {f}"
    assert orc.s.outlet == m["outlet"]
    sc = orc.scalars()
    for k, v in m["scalars"].items():
        assert sc[k] == v or (isinstance(v, float) and np.isnan(v) and np.isnan(sc[k])), f"{name}: {k}"


@pytest.mark.parametrize("name", sorted(MANIFEST["big"]))
def test_oracle_reproduces_512_digests(name):
    m = MANIFEST["big"][name]
    d = synth.fbm_512()
    assert digest(d.z) == m["input_dem"] and digest(d.rain) == m["input_rain"], "generator changed"
    orc = O.Oracle(d.z, d.rain, O.default_params(512, 512, m["flowtype"]), soilT_mode="nonneg")
    for k in range(m["steps"]):
        orc.step(-1000 + k)
    assert orc.s.outlet == m["outlet"]
    for f, h in m["outputs"].items():
        assert digest(orc[f]) == h, f"{name}: {f}"

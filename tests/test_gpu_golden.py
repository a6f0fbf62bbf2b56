"""The HIP path (FP64 form, through the C-ABI) against the golden fixtures
of the oracle (tests/golden/): every stored field bit for bit, and the
digests of the 512^2 runs."""
import hashlib
import json
import os

import numpy as np
import pytest

from parallem_amd import lemhip, synth

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLD, "manifest.json")) as fh:
    MANIFEST = json.load(fh)
FIELDS_F = ["dem", "soilT", "stone", "fines", "fa", "depo", "soilM", "TotB"]
FIELDS_I = ["fd", "SFD", "mask"]


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _run(z, rain, flowtype, mode, runiniter, steps):
    H, W = z.shape
    ctx = lemhip.LemContext(lemhip.default_params(W, H, flowtype, runiniter=runiniter), lemhip.FP64, 0)
    ctx.init_state(z, rain, soilT_mode=mode)
    for k in range(steps):
        ctx.step(-1000 + k)
    return ctx


@pytest.mark.parametrize("name", sorted(MANIFEST["small"]))
def test_gpu_matches_small_fixture(name):
    m = MANIFEST["small"][name]
    with np.load(os.path.join(GOLD, f"{name}.npz")) as g:
        g = {k: g[k] for k in g.files}
    ctx = _run(g["in_dem"], g["in_rain"], m["flowtype"], m["soilT_mode"], m["runiniter"], m["steps"])
    for f in FIELDS_F + FIELDS_I:
        got = ctx.download(f).astype(g[f].dtype)
        assert np.array_equal(got, g[f], equal_nan=True), f"{name}: {f}"
    assert ctx.outlet == m["outlet"]
    s = ctx.summary()
    for k in ("sinkcount", "plateau_rounds1", "plateau_rounds2", "fa_uncomputed", "sed_uncomputed"):
        assert s[k] == m["scalars"][k], f"{name}: {k}"
    ctx.close()


@pytest.mark.parametrize("name", sorted(MANIFEST["big"]))
def test_gpu_matches_512_digests(name):
    m = MANIFEST["big"][name]
    d = synth.fbm_512()
    assert digest(d.z) == m["input_dem"], "generator changed"
    ctx = _run(d.z, d.rain, m["flowtype"], "nonneg", 10, m["steps"])
    assert ctx.outlet == m["outlet"]
    for f, h in m["outputs"].items():
        got = ctx.download(f)
        if f in FIELDS_I:
            got = got.astype(np.int32)
        assert digest(got) == h, f"{name}: {f}"
    ctx.close()

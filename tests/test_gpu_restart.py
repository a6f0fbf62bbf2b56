"""Checkpoint / restart (retriveProcessMatrices, ioGDAL.cpp:662-804): a run
resumed from a snapshot continues bit for bit like the unbroken run, in both
precision forms, in single-DEM and atlas contexts, and the FP64 run stays
bit-identical with the oracle across the break."""
import numpy as np
import pytest

from parallem_amd import lemhip, synth

pytestmark = pytest.mark.gpu

FIELDS = ["dem", "soilT", "fa", "stone", "fines", "soilM", "nut", "soilB", "TotB", "depo", "wC", "wP",
          "Slopes", "fd", "SFD", "mask"]


def _fields(ctx):
    return {f: ctx.download(f) for f in FIELDS}


def _same(a, b, where):
    for f in FIELDS:
        assert np.array_equal(a[f], b[f], equal_nan=True), f"{where}: {f} differs"


@pytest.mark.parametrize("precision", [lemhip.FP64, lemhip.FP32_KAHAN])
def test_resumed_run_equals_unbroken_run(precision):
    dem = synth.fbm_512(side=160, seed_offset=31)
    H, W = dem.z.shape
    p = lemhip.default_params(W, H, 0, runiniter=2)
    a = lemhip.LemContext(p, precision, 0)
    a.init_state(dem.z, dem.rain, soilT_mode="nonneg")
    for k in range(3):
        a.step(-1000 + k)
    snap = a.snapshot()
    for k in range(3, 6):
        a.step(-1000 + k)
    b = lemhip.LemContext(p, precision, 0)
    b.restore(snap)
    for k in range(3, 6):
        b.step(-1000 + k)
    _same(_fields(a), _fields(b), "resumed")
    sa, sb = a.summary(), b.summary()
    for k in ("FA_max", "totE", "totI", "totD", "totweatherC", "totbio", "totwater", "outletcellidx", "sinkcount"):
        assert sa[k] == sb[k], k
    a.close()
    b.close()


def test_resumed_fp64_run_matches_the_oracle():
    from oracle import oracle as O
    dem = synth.fbm_512(side=128, seed_offset=32)
    H, W = dem.z.shape
    a = lemhip.LemContext(lemhip.default_params(W, H, 1), lemhip.FP64, 0)
    a.init_state(dem.z, dem.rain, soilT_mode="nonneg")
    orc = O.Oracle(dem.z, dem.rain, O.default_params(W, H, 1), soilT_mode="nonneg")
    for k in range(2):
        a.step(-1000 + k)
        orc.step(-1000 + k)
    snap = a.snapshot()
    a.close()
    b = lemhip.LemContext(lemhip.default_params(W, H, 1), lemhip.FP64, 0)
    b.restore(snap)
    for k in range(2, 4):
        b.step(-1000 + k)
        orc.step(-1000 + k)
    for f in ("dem", "soilT", "fa", "stone", "TotB", "soilM"):
        assert np.array_equal(b.download(f), orc[f], equal_nan=True), f
    assert np.array_equal(b.download("Slopes"), orc["Slopes"])
    assert np.array_equal(b.download("fd"), orc["fd"])
    assert b.outlet == orc.s.outlet
    b.close()


def test_slot_snapshot_moves_between_atlases():
    d = synth.make_dem(90, 41, pits=3, plateaus=2, outline=True)
    e = synth.make_dem(70, 42, pits=2, plateaus=1)
    p = lemhip.default_params(256, 256, 0)
    a1 = lemhip.LemAtlas(p, 256, 256, lemhip.FP32_KAHAN, 0)
    s1 = a1.add(d.z, d.rain, soilT_mode="nonneg")
    for _ in range(2):
        a1.step()
    snap = a1.snapshot(s1)
    for _ in range(2):
        a1.step()
    a2 = lemhip.LemAtlas(p, 256, 256, lemhip.FP32_KAHAN, 0)
    a2.add(e.z, e.rain, soilT_mode="nonneg")          # another DEM first: a different rectangle
    s2 = a2.open(d.z.shape[1], d.z.shape[0])
    a2.restore(s2, snap, iteration=-998)
    assert a2.rect(s2)[:2] != a1.rect(s1)[:2]
    for _ in range(2):
        a2.step()
    for f in FIELDS:
        assert np.array_equal(a1.download(s1, f), a2.download(s2, f), equal_nan=True), f
    # mismatched size or precision form is refused
    a3 = lemhip.LemAtlas(p, 256, 256, lemhip.FP64, 0)
    s3 = a3.open(d.z.shape[1], d.z.shape[0])
    with pytest.raises(lemhip.LemError):
        a3.restore(s3, snap, iteration=-998)
    s4 = a2.open(40, 40)
    with pytest.raises(lemhip.LemError):
        a2.restore(s4, snap, iteration=-998)
    for a in (a1, a2, a3):
        a.close()

"""Atlas contexts: several DEMs of different sizes stepped together in one
device grid.  Each DEM's fields and totals after every step must equal, bit
for bit, those of a single-DEM context stepped alone on the same inputs —
including a DEM that joins late (its own first iteration), one that leaves,
and one that reuses the freed rectangle.  Both precision forms are checked:
the FP32 form is deterministic too, so alone == in-mix holds exactly there as
well.  (Single-DEM contexts are checked against the oracle in
test_gpu_parity.py.)
"""
import ctypes

import numpy as np
import pytest

from parallem_amd import lemhip, synth

pytestmark = pytest.mark.gpu

FLOAT_FIELDS = ["dem", "soilT", "fa", "runoff", "SlopePtr", "ero", "inci", "geli", "depo", "dz", "wC", "wP",
                "stone", "fines", "soilM", "nut", "soilB", "TotB", "rain", "temp", "Slopes"]
INT_FIELDS = ["fd", "SFD", "mask"]
SUMMARY_KEYS = ("FA_max", "totE", "totI", "totG", "totD", "totweatherC", "totweatherP", "totbio",
                "outletcellidx", "sinkcount", "sinks_pass1")


def _dems():
    a = synth.make_dem(96, 11, pits=3, plateaus=2, height=70)
    b = synth.make_dem(130, 12, pits=5, plateaus=3, outline=True, rain="lognormal")
    c = synth.make_dem(48, 13, pits=1, plateaus=1, height=150)
    rng = np.random.default_rng(3)
    z = rng.uniform(0, 40, (33, 41))   # valid cells on the DEM edge (no nodata frame)
    z[16, :] = 20.0
    d = synth.Dem(z=z, rain=np.full(z.shape, 900.0), seed=0)
    e = synth.make_dem(100, 14, pits=4, plateaus=2, outline=True)
    return {"a": a, "b": b, "c": c, "d": d, "e": e}


class Alone:
    def __init__(self, dem, params_fn, precision):
        H, W = dem.z.shape
        self.ctx = lemhip.LemContext(params_fn(W, H), precision, 0)
        self.ctx.init_state(dem.z, dem.rain, soilT_mode="nonneg")
        self.it = -1000

    def step(self):
        self.ctx.step(self.it)
        self.it += 1


def _compare(atlas, slot, alone, where):
    for f in INT_FIELDS:
        g, o = atlas.download(slot, f), alone.ctx.download(f)
        assert np.array_equal(g, o), f"{where}: {f} differs in {int(np.sum(g != o))} cells"
    for f in FLOAT_FIELDS:
        g, o = atlas.download(slot, f), alone.ctx.download(f)
        assert np.array_equal(g, o, equal_nan=True), f"{where}: {f} differs"
    assert atlas.outlet(slot) == alone.ctx.outlet, f"{where}: outlet"
    s, o = atlas.summary(slot), alone.ctx.summary()
    for k in SUMMARY_KEYS:
        assert s[k] == o[k] or (s[k] != s[k] and o[k] != o[k]), f"{where}: {k} {s[k]} vs {o[k]}"


@pytest.mark.parametrize("precision,flowtype", [(lemhip.FP64, 0), (lemhip.FP64, 1), (lemhip.FP32_KAHAN, 0)])
def test_alone_equals_in_mix(precision, flowtype):
    dems = _dems()

    def params(W, H):
        # runiniter 0: iteration -1000 is each DEM's first, so the outlet rule fires per DEM
        return lemhip.default_params(W, H, flowtype, runiniter=0)

    atlas = lemhip.LemAtlas(params(400, 480), 400, 480, precision, 0)
    alone, slots = {}, {}
    for k in ("a", "b", "c"):
        slots[k] = atlas.add(dems[k].z, dems[k].rain, soilT_mode="nonneg")
        assert slots[k] is not None
        alone[k] = Alone(dems[k], params, precision)
    for k in slots:
        _compare(atlas, slots[k], alone[k], f"init {k}")

    def step(tag):
        atlas.step()
        for k in slots:
            alone[k].step()
        for k in slots:
            _compare(atlas, slots[k], alone[k], f"{tag} {k}")

    step("step 0")
    step("step 1")
    slots["d"] = atlas.add(dems["d"].z, dems["d"].rain, soilT_mode="nonneg")   # joins two steps late
    alone["d"] = Alone(dems["d"], params, precision)
    step("step 2")
    atlas.remove(slots.pop("b"))
    alone.pop("b").ctx.close()
    slots["e"] = atlas.add(dems["e"].z, dems["e"].rain, soilT_mode="nonneg")    # reuses b's rows
    alone["e"] = Alone(dems["e"], params, precision)
    step("step 3")
    step("step 4")
    x0, y0, w, h = atlas.rect(slots["e"])
    assert (w, h) == (100, 100)
    atlas.close()


def test_atlas_room_and_state_errors():
    p = lemhip.default_params(64, 80, 0)
    atlas = lemhip.LemAtlas(p, 64, 80, lemhip.FP64, 0)
    assert atlas.open(64, 10) is None          # needs one column of padding
    s = atlas.open(63, 40)
    assert s is not None
    assert atlas.open(63, 30) is None          # shelves start at multiples of 32: 64 + 31 rows > 80
    t = atlas.open(30, 10)
    assert t is not None and atlas.rect(t)[:2] == (0, 64)
    atlas.remove(s)
    assert atlas.open(63, 40) is not None      # freed rows are reused
    summ = lemhip.LemSummary()
    assert atlas.L.lemhip_summary(atlas.h, ctypes.byref(summ)) == -5      # single-DEM call on an atlas
    assert atlas.L.lemhip_slot_summary(atlas.h, 200, ctypes.byref(summ)) == -1
    atlas.close()


@pytest.mark.parametrize("precision", [lemhip.FP64, lemhip.FP32_KAHAN])
def test_closing_top_shelf_keeps_other_slots(precision):
    """Closing the DEM on the top shelf shrinks the band of rows a step
    covers.  The other DEM (low ground: soil thickness 0, so weathering takes
    the thin-soil calcite term) must keep stepping exactly as alone, with no
    initialisation or climate call in between to refresh cached terms."""
    low = synth.make_dem(96, 21, pits=3, plateaus=2, height=70)      # 96 x 70 cells
    top = synth.make_dem(120, 22, pits=2, plateaus=1, height=150)    # too wide for the first shelf's rest

    def params(W, H):
        return lemhip.default_params(W, H, 0)

    atlas = lemhip.LemAtlas(params(200, 300), 200, 300, precision, 0)
    sa = atlas.add(low.z, low.rain, soilT_mode="nonneg")
    sb = atlas.add(top.z, top.rain, soilT_mode="nonneg")
    assert atlas.rect(sa)[1] == 0 and atlas.rect(sb)[1] == 96   # separate shelves (rows aligned to 32), sb on top
    valid = low.z != -9999.0
    assert np.mean(low.z[valid] < 428.0) > 0.2                   # thin-soil cells (formula soil <= 0.1 m)
    alone = Alone(low, params, precision)
    for k in range(2):
        atlas.step()
        alone.step()
    _compare(atlas, sa, alone, f"before close, step {k}")
    atlas.remove(sb)
    for k in range(2, 5):
        atlas.step()
        alone.step()
        _compare(atlas, sa, alone, f"after closing the top shelf, step {k}")
    assert alone.ctx.summary()["totweatherC"] > 0
    atlas.close()

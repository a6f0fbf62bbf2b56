"""HIP-graph replay of the erosion phase's straight-line kernel sequences
(lem_tuning.use_graphs): every replay, with the step's own eager kernels,
host waits and copies between replays, gives fields and totals bit-identical
to direct launches — in a single-DEM context, and in an atlas whose band of
rows and slot count change between steps (each change recaptures)."""
import numpy as np
import pytest

from parallem_amd import lemhip, synth

pytestmark = pytest.mark.gpu

FIELDS = ["dem", "soilT", "fa", "stone", "fines", "soilM", "nut", "soilB", "TotB", "ero", "inci", "geli", "depo",
          "dz", "wC", "wP", "SlopePtr", "fd"]
TOTALS = ("FA_max", "totE", "totI", "totG", "totD", "totweatherC", "totweatherP", "totbio", "outletcellidx")


@pytest.mark.parametrize("precision", [lemhip.FP64, lemhip.FP32_KAHAN])
def test_graph_replays_equal_direct_launches(precision):
    dem = synth.fbm_512(side=200, seed_offset=71)
    H, W = dem.z.shape
    ctx = []
    for graphs in (0, 1):
        c = lemhip.LemContext(lemhip.default_params(W, H, 0), precision, 0)
        c.init_state(dem.z, dem.rain, soilT_mode="nonneg")
        c.set_tuning(use_graphs=graphs)
        ctx.append(c)
    assert ctx[1].tuning()["use_graphs"] == 1
    for k in range(5):
        for c in ctx:
            c.step(-1000 + k)
        for f in FIELDS:
            assert np.array_equal(ctx[0].download(f), ctx[1].download(f), equal_nan=True), f"step {k}: {f}"
        a, b = ctx[0].summary(), ctx[1].summary()
        for t in TOTALS:
            assert a[t] == b[t], f"step {k}: {t}"
        if k == 2:   # a climate change between replays (temperature-dependent terms recomputed outside the graph)
            for c in ctx:
                c.set_climate(900.0, 12.5)


def test_graph_replays_follow_atlas_changes():
    dems = [synth.make_dem(120, 81, pits=3, plateaus=2), synth.make_dem(90, 82, pits=2, plateaus=1, outline=True),
            synth.make_dem(150, 83, pits=4, plateaus=2)]
    at = []
    for graphs in (0, 1):
        a = lemhip.LemAtlas(lemhip.default_params(300, 400, 0), 300, 400, lemhip.FP32_KAHAN, 0)
        a.set_tuning(use_graphs=graphs)
        at.append(a)
    slots = [[a.add(d.z, d.rain, soilT_mode="nonneg") for d in dems[:2]] for a in at]
    for k in range(6):
        if k == 2:   # a third DEM joins: a new shelf, a taller band
            for a, sl in zip(at, slots):
                sl.append(a.add(dems[2].z, dems[2].rain, soilT_mode="nonneg"))
        if k == 4:   # the first leaves
            for a, sl in zip(at, slots):
                a.remove(sl[0])
        for a in at:
            a.step()
        live = [j for j in range(len(slots[0])) if not (k >= 4 and j == 0)]
        for j in live:
            for f in ("dem", "fa", "soilT", "stone", "wC"):
                assert np.array_equal(at[0].download(slots[0][j], f), at[1].download(slots[1][j], f),
                                      equal_nan=True), f"step {k} slot {j}: {f}"
            s0, s1 = at[0].summary(slots[0][j]), at[1].summary(slots[1][j])
            for t in TOTALS:
                assert s0[t] == s1[t], f"step {k} slot {j}: {t}"
    for a in at:
        a.close()

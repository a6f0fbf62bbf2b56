"""GPU parity: the HIP path through the C-ABI against the CPU oracle.

FP64 form: integer fields (fd, SFD, mask, outlet) and every FP64 field must be
bit-identical with the oracle (shared math routines, no contraction); per-step
totals are sums in a different order and are compared at 1e-12 relative.
"""
import numpy as np
import pytest

from parallem_amd import lemhip, synth

pytestmark = pytest.mark.gpu

FIELDS_F64 = ["dem", "soilT", "soilM", "stone", "fines", "nut", "soilB", "TotB", "fa", "runoff",
              "SlopePtr", "ero", "inci", "geli", "depo", "dz", "wC", "wP"]
ORACLE_NAME = {"runoff": "w"}


def _pair(dem, flowtype=0, soilT_mode="nonneg", precision=lemhip.FP64, runiniter=10):
    from oracle import oracle as O
    H, W = dem.z.shape
    ctx = lemhip.LemContext(lemhip.default_params(W, H, flowtype, runiniter=runiniter), precision, 0)
    ctx.init_state(dem.z, dem.rain, soilT_mode=soilT_mode)
    orc = O.Oracle(dem.z, dem.rain, O.default_params(W, H, flowtype, runiniter=runiniter), soilT_mode=soilT_mode)
    return ctx, orc


def _same(a, b):
    """bit equality with NaN == NaN"""
    return np.array_equal(a, b, equal_nan=True)


def _assert_fields(ctx, orc, where, fields=FIELDS_F64, ints=("fd", "SFD", "mask")):
    for f in ints:
        g, o = ctx.download(f), orc[f]
        assert np.array_equal(g, o), f"{where}: {f} differs in {int(np.sum(g != o))} cells"
    for f in fields:
        g, o = ctx.download(f), orc[ORACLE_NAME.get(f, f)]
        if not _same(g, o):
            d = np.abs(g - o)
            raise AssertionError(f"{where}: {f} differs in {int(np.sum(~((g == o) | (np.isnan(g) & np.isnan(o)))))} "
                                 f"cells, max |diff| {np.nanmax(d):.3e}")
    g = ctx.download("Slopes")
    assert _same(g, orc["Slopes"]), f"{where}: Slopes differ"
    assert ctx.outlet == orc.s.outlet, f"{where}: outlet {ctx.outlet} vs {orc.s.outlet}"


def _assert_summary(ctx, orc, where):
    s = ctx.summary()
    o = orc.scalars()
    for k in ("sinkcount", "fa_uncomputed", "sed_uncomputed", "plateau_rounds1", "plateau_rounds2"):
        assert s[k] == o[k], f"{where}: {k} {s[k]} vs {o[k]}"
    assert s["FA_max"] == o["FA_max"], f"{where}: FA_max"
    for k, ok in (("totwater", "totwater"), ("totE", "totE"), ("totI", "totI"), ("totG", "totG"), ("totD", "totD"),
                  ("totweatherC", "totwC"), ("totweatherP", "totwP"), ("totbio", "totbio")):
        a, b = s[k], o[ok]
        assert abs(a - b) <= 1e-9 * max(1.0, abs(b)), f"{where}: {k} {a} vs {b}"


def test_init_state_matches_oracle():
    dem = synth.fbm_512(side=64)
    ctx, orc = _pair(dem)
    # before the first FlowDirs the reference's fd/SFD are 0 everywhere (Data.cpp:133-155)
    _assert_fields(ctx, orc, "init", ints=("mask",),
                   fields=["dem", "soilT", "soilM", "stone", "fines", "nut", "soilB",
                           "TotB", "rain", "temp", "last_rain", "last_temp"])


@pytest.mark.parametrize("flowtype", [0, 1])
def test_phases_bit_exact_128(flowtype):
    dem = synth.fbm_512(side=128, seed_offset=7)
    ctx, orc = _pair(dem, flowtype)
    it = -1000
    ctx.flow_direction(it); orc.L.orc_flow_direction(orc.p, orc.s, it)
    _assert_fields(ctx, orc, "after FD", fields=["dem"])
    ctx.runoff(); orc.L.orc_runoff(orc.p, orc.s)
    _assert_fields(ctx, orc, "after runoff", fields=["runoff", "soilM"])
    ctx.flow_accum(it); orc.L.orc_flow_accum(orc.p, orc.s, it)
    _assert_fields(ctx, orc, "after FA", fields=["fa"])
    ctx.erosion(it); orc.L.orc_erosion(orc.p, orc.s, it)
    _assert_fields(ctx, orc, "after MOD")
    _assert_summary(ctx, orc, "step 0")


@pytest.mark.parametrize("flowtype", [0, 1])
def test_fbm_512_ten_steps(flowtype):
    """The 512^2 fBm DEM (40 pits, 10 plateaus, 5 % frame, 1 m/yr rain) for 10 steps."""
    dem = synth.fbm_512()
    ctx, orc = _pair(dem, flowtype)
    for k in range(10):
        ctx.step(-1000 + k)
        orc.step(-1000 + k)
        _assert_fields(ctx, orc, f"step {k}", fields=["dem", "fa", "soilT", "stone", "depo"])
        _assert_summary(ctx, orc, f"step {k}")
    _assert_fields(ctx, orc, "final")


def test_reference_soilT_formula_nan_path():
    """Literal createSoilTfromformula: negative soil thickness below 400 m sends
    conc_erosion through 0/0 and NaN; the NaN pattern must match too."""
    dem = synth.fbm_512(side=96, seed_offset=3)
    ctx, orc = _pair(dem, soilT_mode="formula")
    # input diagnostics: valid cells with 0 < z < 400 (negative formula soil
    # thickness), and valid cells in columns 0..P-1 of rows 1.. (P = 16 - w % 16)
    z = dem.z
    valid = z != -9999.0
    s = ctx.summary()
    assert s["soilT_negative"] == int((valid & (z > 0) & (z < 400)).sum()) > 0
    w = z.shape[1]
    P = min((w // 16 + 1) * 16 - w, w)
    assert s["phantom_cells"] == int(valid[1:, :P].sum())
    for k in range(3):
        ctx.step(-1000 + k)
        orc.step(-1000 + k)
        _assert_fields(ctx, orc, f"formula step {k}", fields=["dem", "fa", "soilT", "inci"])
    assert np.isnan(orc["inci"]).any()


def test_outlet_first_iteration_rule():
    """runiniter = 0 makes iteration -1000 the first: the outlet moves to the
    first cell holding the maximum flow accumulation (mfd_simple.cu:336-349),
    and the surface update then lifts every cell at or below the new base
    level (updates.cu:151-161)."""
    dem = synth.fbm_512(side=96, seed_offset=5)
    ctx, orc = _pair(dem, runiniter=0)
    o_init = orc.s.outlet
    for k in range(2):
        ctx.step(-1000 + k)
        orc.step(-1000 + k)
        assert ctx.outlet == orc.s.outlet
        _assert_fields(ctx, orc, f"runin0 step {k}")
        _assert_summary(ctx, orc, f"runin0 step {k}")
    assert orc.s.outlet != o_init


def test_outlet_check_lowers_a_raised_outlet():
    """The outlet check (routeflow.cu:171-184, 739-741): when the outlet is
    not the lowest valid cell at the start of a step, it is set to
    lowest - 1e-4 after FlowDirs has read the old surface and before flow
    accumulation.  A step always ends with the outlet as the strict minimum
    (the surface update lifts every cell at or below the base level), so the
    branch fires only after the caller moves the outlet or edits the surface;
    here the outlet is moved to a higher cell between steps."""
    dem = synth.fbm_512(side=96, seed_offset=5)
    ctx, orc = _pair(dem)
    ctx.step(-1000)
    orc.step(-1000)
    valid = np.flatnonzero(orc["mask"] == 1)
    for k, pick in ((1, valid[len(valid) // 3]), (2, valid[2 * len(valid) // 3])):
        ctx.outlet = int(pick)
        orc.s.outlet = int(pick)
        z = orc["dem"].copy()
        lowest = min(1000.0, float(z[valid].min()))
        assert lowest < z[pick]
        ctx.step(-1000 + k)
        orc.step(-1000 + k)
        _assert_fields(ctx, orc, f"moved outlet, step {k}")
        _assert_summary(ctx, orc, f"moved outlet, step {k}")
        # lowered to lowest - 1e-4, then the surface update pinned it
        # 0.00012 below the base level it defined (updates.cu:151,165-174)
        assert orc["dem"][pick] == ((lowest - 1e-4) - 0.00012) - 0.00012


def test_no_nodata_frame_and_tiny_grids():
    """Grids whose edge cells are valid (out-of-grid neighbours act as nodata)."""
    rng = np.random.default_rng(11)
    for side in (3, 17, 65):
        z = rng.uniform(0, 50, (side, side))
        z[side // 2, :] = 25.0  # a flat ridge
        dem = synth.Dem(z=z, rain=np.full(z.shape, 1000.0), seed=0)
        for ft in (0, 1):
            ctx, orc = _pair(dem, ft)
            for k in range(2):
                ctx.step(-1000 + k)
                orc.step(-1000 + k)
            _assert_fields(ctx, orc, f"side {side} ft {ft}")


def test_batch_of_catchment_outlines():
    for d in synth.catchment_batch(3, seed=99, lo=64, hi=160):
        ctx, orc = _pair(d)
        for k in range(2):
            ctx.step(-1000 + k)
            orc.step(-1000 + k)
        _assert_fields(ctx, orc, f"outline {d.shape}")


def test_gpu_math_bit_identical_to_host():
    from oracle import oracle as O
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(-50, 50, 4000), rng.uniform(-745, 709, 2000), [0.0, -0.0, 1e-300, 700.0]])
    pos = np.concatenate([rng.uniform(1e-12, 1e4, 4000), np.exp(rng.uniform(-700, 700, 2000)), [1.0, 10.0, 0.5, 2.0]])
    y = rng.uniform(-3, 3, x.size)
    for fn, arg, arg2 in ((0, x, None), (1, pos, None), (2, pos, None), (3, pos, y), (4, x, None), (5, x, None)):
        g = lemhip.math_eval(fn, arg, arg2)
        h = O.math_eval(fn, arg, arg2)
        assert _same(g, h), f"math fn {fn} differs"


def test_temperature_change_mid_run():
    """MOD part 2 keeps per-cell terms computed from the temperatures; a new
    temperature field (upload) or a new climate (set_climate) between steps
    must be picked up, bit for bit with the oracle given the same change."""
    from oracle import oracle as O
    dem = synth.fbm_512(side=128, seed_offset=11)
    ctx, orc = _pair(dem)
    it = -1000
    for k in range(2):
        ctx.step(it + k); orc.step(it + k)
    # a warmer, uneven field through an upload
    t = orc["temp"].copy()
    valid = orc["mask"] == 1
    t[valid] += 3.0 + 0.01 * (np.arange(t.size)[valid] % 97)
    orc["temp"][:] = t
    ctx.upload("temp", t.reshape(dem.z.shape))
    for k in range(2, 4):
        ctx.step(it + k); orc.step(it + k)
        _assert_fields(ctx, orc, f"after upload, step {k}", fields=["dem", "soilT", "stone", "wC", "wP"])
    # a new climate through the setter (the oracle's own climate routine)
    ctx.set_climate(1.5, 12.0)
    p = O.default_params(*dem.z.shape[::-1])
    orc.L.orc_SetClimateMatrices(O.C.byref(p), O.iptr(orc["mask"]), O.dptr(orc["dem"]), 1.5, 12.0,
                                 O.dptr(orc["rain"]), O.dptr(orc["temp"]))
    _assert_fields(ctx, orc, "after set_climate", ints=("mask",), fields=["rain", "temp"])
    for k in range(4, 6):
        ctx.step(it + k); orc.step(it + k)
        _assert_fields(ctx, orc, f"after set_climate, step {k}", fields=["dem", "soilT", "stone", "wC", "wP"])
    _assert_summary(ctx, orc, "final")


@pytest.mark.parametrize("flowtype", [0, 1])
def test_long_chain_of_basins(flowtype):
    """A valley of ~100 one-cell pits whose sills rise away from the grid
    edge, with stretches of equal sills: each basin's lightest spill leads
    to its downstream neighbour, so the spill heights come from chains of
    basins joining in one phase (many pointer-jumping passes) and from tied
    weights.  Directions, raised surface and everything downstream match
    the oracle's literal Kruskal pass bit for bit over three steps."""
    H, W = 25, 640
    y = np.arange(H)[:, None].astype(np.float64)
    x = np.arange(W)[None, :].astype(np.float64)
    floor = 100.0 + 0.05 * np.floor(x / 40.0) * 40.0 + 0.01 * x * ((x // 80) % 2)   # rising, with level stretches
    z = floor + 2.0 * np.abs(y - 12.0)
    for px in range(8, W - 4, 6):
        z[12, px] -= 3.0
    z[0, :] = -9999.0
    z[-1, :] = -9999.0
    z[:, -1] = -9999.0
    dem = synth.Dem(z=z, rain=np.full(z.shape, 1000.0), seed=0)
    ctx, orc = _pair(dem, flowtype)
    for k in range(3):
        ctx.step(-1000 + k)
        orc.step(-1000 + k)
        _assert_fields(ctx, orc, f"step {k}", fields=["dem", "fa", "soilT", "stone", "depo"])
        _assert_summary(ctx, orc, f"step {k}")
    assert ctx.summary()["nlabels"] > 90


def _wide_flats():
    """A 200 x 262 surface falling west (and gently toward its middle row)
    with a flat mesa (101 x 181 cells at 400 m) behind a 450 m wall, drained
    through one notch on its west side, and a basin (51 x 221 cells,
    200-210 m) behind a 260 m rim that spills east over one 230 m sill.
    (A plane falling toward a corner of the frame would make the MFD
    incision NaN next to the frame, in the oracle and here alike.)"""
    H, W = 200, 262
    y = np.arange(H)[:, None].astype(np.float64)
    x = np.arange(W)[None, :].astype(np.float64)
    z = 100.0 + 0.5 * x + 0.01 * np.abs(y - 100.0)
    z[19:122, 59:242] = 450.0
    z[20:121, 60:241] = 400.0
    z[70, 59] = 400.0
    rng = np.random.Generator(np.random.PCG64(7))
    z[134:187, 29:252] = 260.0
    z[135:186, 30:251] = 200.0 + 10.0 * rng.random((51, 221))
    z[160, 251] = 230.0
    z[0, :] = z[-1, :] = -9999.0
    z[:, 0] = z[:, -1] = -9999.0
    return synth.Dem(z=z, rain=np.full(z.shape, 1000.0), seed=0)


@pytest.mark.parametrize("flowtype", [0, 1])
def test_flats_drained_through_one_cell(flowtype):
    """Plateau fronts that start at a single exit cell and cross many
    plateau tiles in every direction: the mesa on the dem (181 rounds in the
    first step) and the flooded basin on the raised surface (221 rounds), so
    the first launch holds only the tiles next to the notch and the sill, and
    each later launch only the tiles within reach of the last round's front.
    Directions, slope marks, round counts and everything downstream match the
    oracle bit for bit, at every rounds-per-launch choice."""
    dem = _wide_flats()
    for k_launch in (4, 8, 16):
        ctx, orc = _pair(dem, flowtype)
        ctx.set_tuning(plateau_rounds_per_launch=k_launch)
        for k in range(2):
            ctx.step(-1000 + k)
            orc.step(-1000 + k)
            _assert_fields(ctx, orc, f"K={k_launch} step {k}", fields=["dem", "fa", "soilT", "stone", "depo"])
            _assert_summary(ctx, orc, f"K={k_launch} step {k}")
        s = ctx.summary()
        assert s["plateau_rounds2"] > 200
        ctx.close()

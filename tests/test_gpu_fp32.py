"""Error bounds of the FP32 + compensated-accumulation form against FP64.

The FP64 reference here is the GPU's plain FP64 form, which is bit-identical
to the CPU oracle (test_gpu_parity.py); at 4096^2 x 100 steps the CPU oracle
itself would take hours.  Bounds (stated, and asserted):

* flow accumulation, one step from identical state (identical directions by
  construction, FD runs on the FP64 dem in both forms): max relative error
  <= 1e-6 over cells whose FP64 value is below 3e38 (MFD path counts can
  overflow FP32 otherwise);
* elevation and flow accumulation over one step from the same state (the
  FP64 trajectory's state after 50 steps of the 4096^2 fBm DEM, rounded into the
  FP32 form): max |dz| <= 1e-4 m, flow accumulation max relative error
  <= 1e-6 where the directions agree;
* the 100-step trajectories of the 4096^2 fBm DEM (SFD) drift apart only
  through discrete flips: a direction at a near-exact slope tie taken
  differently after ~1e-8 m differences in the surface.  Asserted: every
  cell off the downstream paths of the flipped cells (and their
  neighbourhoods) within 1e-4 m after 100 steps, every cell above 1e-4 m on
  such a path, at most 1e-6 of the cells with a different direction at any
  checked step; the maximum and the counts are printed;
* the FP32 form against the CPU oracle itself (FP64) on the 512^2 DEM for 10
  steps, a 128^2 MFD DEM for 3 steps, a 256^2 MFD DEM (flow past the float
  range) and a 1024^2 DEM for one step:
  fa within 1e-6 relative (where the FP64 value is below 3e38) and elevation
  within 1e-4 m after every step, directions and cycle cells equal.
"""
import numpy as np
import pytest

from parallem_amd import lemhip, synth

pytestmark = pytest.mark.gpu


def _pair(dem, flowtype):
    H, W = dem.z.shape
    out = []
    for prec in (lemhip.FP64, lemhip.FP32_KAHAN):
        c = lemhip.LemContext(lemhip.default_params(W, H, flowtype), prec, 0)
        c.init_state(dem.z, dem.rain, soilT_mode="nonneg")
        out.append(c)
    return out


def _fa_rel_error(c64, c32, it):
    for c in (c64, c32):
        c.flow_direction(it)
        c.runoff()
        c.flow_accum(it)
    fd64, fd32 = c64.download("fd"), c32.download("fd")
    assert np.array_equal(fd64, fd32), "directions differ on an identical dem"
    a, b = c64.download("fa"), c32.download("fa")
    m = (c64.download("mask") == 1) & (np.abs(a) < 3e38) & (a > 0)
    return float(np.max(np.abs(b[m] - a[m]) / a[m]))


@pytest.mark.parametrize("flowtype,side", [(0, 4096), (1, 1024)])
def test_flow_accumulation_relative_error(flowtype, side):
    dem = synth.fbm_4096(side=side)
    c64, c32 = _pair(dem, flowtype)
    err = _fa_rel_error(c64, c32, -1000)
    print(f"\n[fp32] flow accumulation max rel error, {side}^2 flowtype {flowtype}: {err:.3e}")
    assert err <= 1e-6


STATE = ["dem", "soilT", "Slopes", "soilM", "stone", "fines", "nut", "soilB", "TotB", "rain", "temp",
         "last_rain", "last_temp"]


def test_one_step_from_the_same_state():
    dem = synth.fbm_4096()
    c64, c32 = _pair(dem, 0)
    for k in range(50):
        c64.step(-1000 + k)
    for f in STATE:
        c32.upload(f, c64.download(f))
    c32.outlet = c64.outlet
    c64.step(-950)
    c32.step(-950)
    m = c64.download("mask") == 1
    dz = float(np.max(np.abs(c64.download("dem")[m] - c32.download("dem")[m])))
    fd_same = c64.download("fd")[m] == c32.download("fd")[m]
    a, b = c64.download("fa")[m], c32.download("fa")[m]
    ok = fd_same & (a > 0) & (a < 3e38)
    rel = float(np.max(np.abs(b[ok] - a[ok]) / a[ok]))
    print(f"\n[fp32] one step from the FP64 state at step 50: max |dz| = {dz:.3e} m, fa max rel = {rel:.3e}, "
          f"direction mismatches = {int(np.sum(~fd_same))}")
    assert dz <= 1e-4
    assert rel <= 1e-6


DIRS = [(-1, 0), (-1, 1), (0, 1), (1, 1), (1, 0), (1, -1), (0, -1), (-1, -1)]   # N..NW as (drow, dcol)
CODE_DIR = {64: 0, 128: 1, 1: 2, 2: 3, 4: 4, 8: 5, 16: 6, 32: 7}


def _downstream(fd, W, H, start, out):
    """Mark the cells on the receiver path from `start` (single flow directions)."""
    i = int(start)
    for _ in range(W * H):
        if out[i]:
            return
        out[i] = True
        d = CODE_DIR.get(int(fd[i]))
        if d is None:
            return
        r, c = divmod(i, W)
        r, c = r + DIRS[d][0], c + DIRS[d][1]
        if not (0 <= r < H and 0 <= c < W):
            return
        i = r * W + c


def test_trajectory_drift_100_steps_fbm_4096():
    """Over 100 steps the two trajectories part only through discrete flips
    (the FP64 surfaces and soil thicknesses differ by ~1e-8 between the
    forms): a direction at a near-exact slope tie, whose upstream flow then
    takes another path, so both downstream paths see other discharge; or a
    branch of the per-cell update at a threshold — the calcite term's soil
    thickness > 0.1 m (chemical weathering jumps by a factor of 10^(1/3)),
    the bare-rock reset at thickness 0 (stone and moisture reset), the
    evapotranspiration branch (moisture drawdown jumps) — whose changed
    moisture also changes the cell's runoff below it.  Asserted: every cell
    off the flipped cells' downstream paths (and their 8-neighbourhoods,
    which see the changed slopes) within 1e-4 m after 100 steps; every cell
    above 1e-4 m on such a path; at most 1e-6 of the cells with a different
    direction at any checked step.  Counts are printed."""
    dem = synth.fbm_4096()
    c64, c32 = _pair(dem, 0)
    H, W = dem.z.shape
    m = c64.download("mask") == 1
    flips = []
    branch = []
    affected = np.zeros(m.shape, bool)
    for k in range(100):
        c64.step(-1000 + k)
        c32.step(-1000 + k)
        f64, f32 = c64.download("fd"), c32.download("fd")
        jumps = m & ((np.abs(c64.download("wC") - c32.download("wC")) > 1e-6)
                     | (np.abs(c64.download("stone") - c32.download("stone")) > 1.0)
                     | (np.abs(c64.download("soilM") - c32.download("soilM")) > 1e-3))
        diff = np.flatnonzero((m & (f64 != f32)) | jumps)
        for x in diff:
            _downstream(f64, W, H, x, affected)
            _downstream(f32, W, H, x, affected)
        branch.append(int(jumps.sum()))
        if k % 10 == 9 or k == 0:
            flips.append((k, int(np.sum(m & (f64 != f32)))))
    grown = affected.reshape(H, W).copy()
    for dr, dc in DIRS:
        grown |= np.roll(np.roll(affected.reshape(H, W), dr, 0), dc, 1)
    grown = grown.reshape(-1)
    d = np.abs(c64.download("dem") - c32.download("dem"))
    d[~m] = 0
    worst, p9999 = float(d.max()), float(np.percentile(d[m], 99.99))
    off = m & ~grown
    worst_off = float(d[off].max())
    above = int(np.sum(d > 1e-4))
    above_on = int(np.sum((d > 1e-4) & grown))
    print(f"\n[fp32] 4096^2 fBm, 100 steps: max |dz| = {worst:.3e} m, 99.99th percentile {p9999:.3e} m; "
          f"cells above 1e-4 m: {above} ({above_on} below a flip); off the flipped paths: max |dz| = "
          f"{worst_off:.3e} m; cells on them: {int(grown.sum())}; direction mismatches (step, cells): {flips}; "
          f"branch flips per step: total {sum(branch)}, max {max(branch)}")
    assert worst_off <= 1e-4
    assert above == above_on
    assert max(f for _, f in flips) <= 1e-6 * m.sum()


def _assert_fp32_near_oracle(dem, flowtype, steps, where):
    """The FP32 form against the CPU oracle (FP64) directly: fd/SFD exact
    while the surfaces agree, flow accumulation within 1e-6 relative and
    elevation within 1e-4 m after every step."""
    from oracle import oracle as O
    H, W = dem.z.shape
    c = lemhip.LemContext(lemhip.default_params(W, H, flowtype), lemhip.FP32_KAHAN, 0)
    c.init_state(dem.z, dem.rain, soilT_mode="nonneg")
    orc = O.Oracle(dem.z, dem.rain, O.default_params(W, H, flowtype), soilT_mode="nonneg")
    m = orc["mask"] == 1
    worst_fa = worst_z = 0.0
    for k in range(steps):
        c.step(-1000 + k)
        orc.step(-1000 + k)
        same = c.download("fd")[m] == orc["fd"][m]
        assert same.all(), f"{where} step {k}: {int(np.sum(~same))} directions differ"
        a, b = orc["fa"][m], c.download("fa")[m]
        ok = (a > 0) & (a < 3e38)
        worst_fa = max(worst_fa, float(np.max(np.abs(b[ok] - a[ok]) / a[ok])))
        worst_z = max(worst_z, float(np.max(np.abs(c.download("dem")[m] - orc["dem"][m]))))
        assert c.summary()["fa_uncomputed"] == orc.scalars()["fa_uncomputed"], f"{where} step {k}: cycle cells"
    print(f"\n[fp32 vs oracle] {where}: fa max rel {worst_fa:.3e}, max |dz| {worst_z:.3e} m over {steps} steps")
    assert worst_fa <= 1e-6
    assert worst_z <= 1e-4
    c.close()


def test_fp32_against_the_oracle_fbm_512_ten_steps():
    _assert_fp32_near_oracle(synth.fbm_512(), 0, 10, "512^2 SFD")


def test_fp32_against_the_oracle_mfd():
    # MFD passes the full flow to every receiver (prop = 1), so fa grows like
    # path counts: 4.5e32 at 128^2 (inside the float range), 2.2e51 at 256^2,
    # where the FP32 form saturates at inf and every consumer clamps it as the
    # FP64 form clamps its huge finite value; fa is compared below 3e38
    # (after one step the two surfaces differ by ~1e-7 m, and a multi-bit MFD
    # code, which takes every lower neighbour, flips at near-equal heights
    # more readily than a single direction: the 256^2 case is checked over
    # the step whose directions come from the identical initial surface)
    _assert_fp32_near_oracle(synth.fbm_512(side=128, seed_offset=4), 1, 3, "128^2 MFD")
    _assert_fp32_near_oracle(synth.fbm_512(side=256, seed_offset=4), 1, 1, "256^2 MFD")


def test_fp32_against_the_oracle_1024_one_step():
    _assert_fp32_near_oracle(synth.fbm_4096(side=1024), 0, 1, "1024^2 SFD")


def test_integer_accumulation_equals_the_engine():
    """FP32 form, single flow directions: the exact integer subtree sums and
    the dependency engine's compensated float sums agree to 1e-9 relative on
    every cell, the outlet cycle included (runiniter = 0 moves the outlet onto
    the largest river, whose west neighbour then drains back into it)."""
    dem = synth.fbm_512(side=640, seed_offset=61)
    H, W = dem.z.shape
    ctx = {}
    for integer in (1, 0):
        c = lemhip.LemContext(lemhip.default_params(W, H, 0, runiniter=0), lemhip.FP32_KAHAN, 0)
        c.init_state(dem.z, dem.rain, soilT_mode="nonneg")
        c.set_tuning(fp32_integer_fa=integer)
        ctx[integer] = c
    for k in range(3):
        for c in ctx.values():
            c.step(-1000 + k)
        a, b = ctx[1].download("fa"), ctx[0].download("fa")
        m = ctx[1].download("mask") == 1
        assert np.array_equal(ctx[1].download("fd"), ctx[0].download("fd"))
        rel = np.abs(a[m] - b[m]) / np.maximum(b[m], 1e-30)
        assert float(rel.max()) <= 1e-9, f"step {k}: {float(rel.max())}"
        assert ctx[1].summary()["fa_uncomputed"] == ctx[0].summary()["fa_uncomputed"]
        assert np.array_equal(ctx[1].download("dem"), ctx[0].download("dem")) or k > 0
    print(f"\n[fp32] integer vs engine accumulation: cycle cells {ctx[1].summary()['fa_uncomputed']}")
    assert ctx[1].tuning()["fp32_integer_fa"] == 1 and ctx[0].tuning()["fp32_integer_fa"] == 0


@pytest.mark.parametrize("rounds", [1, 3])
def test_tree_sediment_equals_the_engine(rounds):
    """FP32 form, single flow directions: sediment routing by the engine's
    first rounds plus a contraction of the rest of the donor forest (the
    default, with three rounds) against the engine's rounds to the end, one erosion phase at a
    time from identical state, the outlet cycle included (runiniter = 0).
    The same cells compute; deposition, soil thickness and every cell whose
    stone value is reached without a composed chain are bit-identical; stone
    values differ by at most 1e-6 relative (a few float ulps)."""
    dem = synth.fbm_512(side=640, seed_offset=63)
    H, W = dem.z.shape
    ctx = {}
    for tree in (1, 0):
        c = lemhip.LemContext(lemhip.default_params(W, H, 0, runiniter=0), lemhip.FP32_KAHAN, 0)
        c.init_state(dem.z, dem.rain, soilT_mode="nonneg")
        c.set_tuning(fp32_tree_sediment=rounds if tree else 0)
        ctx[tree] = c
    m = ctx[0].download("mask") == 1
    for k in range(3):
        if k:
            for f in STATE:
                ctx[1].upload(f, ctx[0].download(f))
            ctx[1].outlet = ctx[0].outlet
        for c in ctx.values():
            c.flow_direction(-1000 + k)
            c.runoff()
            c.flow_accum(-1000 + k)
            c.erosion(-1000 + k)
        t, e = ctx[1], ctx[0]
        assert t.summary()["sed_uncomputed"] == e.summary()["sed_uncomputed"]
        for f in ("depo", "soilT", "ero", "inci", "geli"):
            assert np.array_equal(t.download(f), e.download(f), equal_nan=True), f"step {k}: {f}"
        a, b = t.download("stone")[m], e.download("stone")[m]
        rel = float(np.max(np.abs(a.astype(np.float64) - b) / np.abs(b)))
        same = float(np.mean(a == b))
        print(f"\n[fp32] tree ({rounds} engine rounds) vs engine sediment, step {k}: stone max rel {rel:.2e}, "
              f"bit-equal {same:.4f}, "
              f"uncomputed {t.summary()['sed_uncomputed']}")
        assert rel <= 1e-6
        fa_, fb_ = t.download("fines")[m], e.download("fines")[m]
        assert float(np.max(np.abs(fa_ - fb_))) <= 1e-4
    assert ctx[1].tuning()["fp32_tree_sediment"] == rounds and ctx[0].tuning()["fp32_tree_sediment"] == 0


def test_integer_accumulation_falls_back_on_negative_weights():
    """A negative rain cell gives a negative runoff weight, which the
    integer sums cannot hold: the step falls back to the engine and still
    matches the FP64 form."""
    dem = synth.fbm_512(side=200, seed_offset=62)
    rain = dem.rain.copy().reshape(-1)
    valid = (dem.z != -9999.0).reshape(-1)
    rain[np.flatnonzero(valid)[500]] = -50.0
    rain = rain.reshape(dem.z.shape)
    H, W = dem.z.shape
    c64, c32 = (lemhip.LemContext(lemhip.default_params(W, H, 0), p, 0) for p in (lemhip.FP64, lemhip.FP32_KAHAN))
    for c in (c64, c32):
        c.init_state(dem.z, rain, soilT_mode="nonneg")
        c.flow_direction(-1000)
        c.runoff()
        c.flow_accum(-1000)
    a, b = c64.download("fa"), c32.download("fa")
    m = valid.reshape(-1) & (np.abs(a) > 1e-12)
    assert float(np.max(np.abs(b[m] - a[m]) / np.abs(a[m]))) <= 1e-6


def test_integer_accumulation_outlet_cycle():
    """A valley whose lowest cell (the outlet) sits mid-row: its west
    neighbour drains east into it while fd[outlet] = WEST sends it back, a
    2-cycle whose cells never complete in the reference (fa = w).  The FP32
    form's integer sums must find the same cycle as the oracle."""
    from oracle import oracle as O
    H, W = 40, 41
    z = np.full((H, W), -9999.0)
    r, c = np.mgrid[0:H, 0:W]
    inner = (r > 0) & (r < H - 1) & (c > 0) & (c < W - 1)
    z[inner] = (10.0 * np.abs(c - 20) + 3.0 * r + 0.01 * ((r * 7 + c * 3) % 5))[inner]
    dem = synth.Dem(z=z, rain=np.full(z.shape, 1000.0), seed=0)
    ctx = lemhip.LemContext(lemhip.default_params(W, H, 0), lemhip.FP32_KAHAN, 0)
    ctx.init_state(dem.z, dem.rain, soilT_mode="nonneg")
    orc = O.Oracle(dem.z, dem.rain, O.default_params(W, H, 0), soilT_mode="nonneg")
    assert orc.s.outlet == 1 * W + 20
    for k in range(2):
        ctx.step(-1000 + k)
        orc.step(-1000 + k)
        assert orc.scalars()["fa_uncomputed"] >= 2
        assert ctx.summary()["fa_uncomputed"] == orc.scalars()["fa_uncomputed"]
        assert ctx.summary()["sed_uncomputed"] == orc.scalars()["sed_uncomputed"]
        a, b = orc["fa"], ctx.download("fa")
        m = orc["mask"] == 1
        assert float(np.max(np.abs(b[m] - a[m]) / a[m])) <= 1e-6
        o = orc.s.outlet
        assert b[o] == np.float32(ctx.download("runoff")[o])      # cycle cells keep fa = w


def test_tree_sediment_falls_back_on_non_finite_terms():
    """The literal soil-thickness formula makes incision NaN below 400 m, so
    deposition and the stone-mixing weights of cells downstream are not
    finite: the sediment pass of the FP32 form then runs the engine's
    rounds to the end (nothing of the state changes before the check), and
    the results equal a context that always does, bit for bit."""
    dem = synth.fbm_512(side=96, seed_offset=3)
    H, W = dem.z.shape
    out = []
    for tree in (3, 0):
        c = lemhip.LemContext(lemhip.default_params(W, H, 0), lemhip.FP32_KAHAN, 0)
        c.init_state(dem.z, dem.rain, soilT_mode="formula")
        c.set_tuning(fp32_tree_sediment=tree)
        for k in range(3):
            c.step(-1000 + k)
            assert c.summary()["sed_open"] == 0, f"step {k}: the contraction ran on non-finite terms"
        out.append({f: c.download(f) for f in ("dem", "stone", "fines", "soilT", "depo", "inci")})
        c.close()
    assert np.isnan(out[0]["inci"]).any()
    for f in out[0]:
        assert np.array_equal(out[0][f], out[1][f], equal_nan=True), f

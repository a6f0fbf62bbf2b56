"""Launch tuning never changes results: every dependency-engine tile shape,
launch batch and plateau rounds-per-launch choice, and an autotuned step,
give bit-identical fields (FP64 form; the oracle comparison of the default
choice is in test_gpu_parity.py)."""
import numpy as np
import pytest

from parallem_amd import lemhip, synth

pytestmark = pytest.mark.gpu

FIELDS = ["dem", "fa", "soilT", "stone", "depo", "Slopes", "fd", "SFD"]
CHOICES = [dict(dag_tile=(32, 32)), dict(dag_tile=(16, 64), dag_batch=4), dict(dag_tile=(128, 8)),
           dict(dag_tile=(64, 32), plateau_rounds_per_launch=4), dict(dag_tile=(32, 64), plateau_rounds_per_launch=16)]


def _run(dem, flowtype, tuning=None, autotune=False, steps=3):
    H, W = dem.z.shape
    ctx = lemhip.LemContext(lemhip.default_params(W, H, flowtype), lemhip.FP64, 0)
    ctx.init_state(dem.z, dem.rain, soilT_mode="nonneg")
    if tuning:
        ctx.set_tuning(**tuning)
    for k in range(steps):
        if autotune and k == 1:
            ctx.autotune()
        ctx.step(-1000 + k)
    out = {f: ctx.download(f) for f in FIELDS}
    out["tuning"] = ctx.tuning()
    ctx.close()
    return out


@pytest.mark.parametrize("flowtype", [0, 1])
def test_every_tuning_gives_identical_results(flowtype):
    dem = synth.fbm_512(side=300, seed_offset=31)
    ref = _run(dem, flowtype)
    for t in CHOICES + [None]:
        got = _run(dem, flowtype, tuning=t, autotune=t is None)
        for f in FIELDS:
            assert np.array_equal(got[f], ref[f], equal_nan=True), f"{t}: {f}"
    tuned = got["tuning"]
    assert tuned["plateau_rounds_per_launch"] in (4, 8, 16)
    assert tuple(tuned["dag_tile"]) in {(64, 16), (32, 32), (16, 64), (128, 8), (64, 32), (32, 64)}
    assert all(x > 0 for x in tuned["tuned_ms"][:6])


def test_bad_tuning_is_rejected():
    ctx = lemhip.LemContext(lemhip.default_params(32, 32), lemhip.FP64, 0)
    with pytest.raises(lemhip.LemError):
        ctx.set_tuning(dag_tile=(48, 48))
    with pytest.raises(lemhip.LemError):
        ctx.set_tuning(plateau_rounds_per_launch=5)
    ctx.close()


@pytest.mark.parametrize("precision", [lemhip.FP64, lemhip.FP32_KAHAN])
def test_autotune_runs_once(precision):
    """lemhip_autotune tunes on the next step only: a shape set afterwards is
    kept by the following steps and the phase-level calls.  (In the FP32
    form the tile shape serves the engine's sediment rounds, so those are
    switched on.)"""
    dem = synth.fbm_512(side=200, seed_offset=33)
    H, W = dem.z.shape
    ctx = lemhip.LemContext(lemhip.default_params(W, H, 0), precision, 0)
    ctx.init_state(dem.z, dem.rain, soilT_mode="nonneg")
    if precision == lemhip.FP32_KAHAN:
        ctx.set_tuning(fp32_tree_sediment=0)
    ctx.autotune()
    ctx.step(-1000)
    first = ctx.tuning()
    assert all(x > 0 for x in first["tuned_ms"][:6])
    other = (16, 64) if tuple(first["dag_tile"]) != (16, 64) else (64, 16)
    ctx.set_tuning(dag_tile=other, plateau_rounds_per_launch=4)
    ctx.step(-999)
    ctx.flow_direction(-998)
    ctx.runoff()
    ctx.flow_accum(-998)
    t = ctx.tuning()
    assert tuple(t["dag_tile"]) == other and t["plateau_rounds_per_launch"] == 4
    ctx.close()


@pytest.mark.parametrize("tree", [0, 1])
def test_fp32_sediment_tuning_restores_the_state(tree):
    """FP32 form with integer flow accumulation: with the engine's sediment
    rounds (tree 0) the autotuner times sediment routing on the step's own
    data, saving and restoring the fields the pass updates in place; with
    the first-round-plus-contraction routing (tree 1, fixed tile shape) it
    times nothing there.  The tuned step must equal an untuned one bit for
    bit."""
    dem = synth.fbm_512(side=300, seed_offset=35)
    H, W = dem.z.shape
    out = []
    for tune in (True, False):
        ctx = lemhip.LemContext(lemhip.default_params(W, H, 0), lemhip.FP32_KAHAN, 0)
        ctx.init_state(dem.z, dem.rain, soilT_mode="nonneg")
        ctx.set_tuning(fp32_tree_sediment=tree)
        if tune:
            ctx.autotune()
        for k in range(2):
            ctx.step(-1000 + k)
        out.append({f: ctx.download(f) for f in FIELDS})
        if tune:
            t = ctx.tuning()
            timed = [x > 0 for x in t["tuned_ms"][:6] + t["tuned_ms"][11:14]]
            assert all(timed) if tree == 0 else not any(timed)
        ctx.close()
    for f in FIELDS:
        assert np.array_equal(out[0][f], out[1][f], equal_nan=True), f

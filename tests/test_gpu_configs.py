"""The three larger workloads at their own sizes:

* batched uneven catchments: 64 fBm catchments with sides log-uniform in
  [256, 2048] and irregular outlines, FP32 form, two atlases stepped together
  for 50 steps; a sample of the DEMs run alone must match their in-mix
  results bit for bit, and every DEM stays finite with non-negative flow;
* serving: requests with DEM sides log-uniform in [256, 4096] and a forcing
  series, continuously batched; each result equals the request run alone;
* one 16384^2 DEM (268 M cells) resident in the FP32 form: its footprint is
  what the library states per cell, well inside the device memory, and two
  steps run with finite results.
"""
import numpy as np
import pytest

from parallem_amd import batchrun, lemhip, serving, synth

pytestmark = pytest.mark.gpu


def test_batched_catchments_64_fifty_steps():
    dems = synth.catchment_batch(64, 20261203, 256, 2048)
    groups = []
    for members, w, h in batchrun.pack(dems, 2, 8192):
        at = lemhip.LemAtlas(lemhip.default_params(w, h, 0), w, h, lemhip.FP32_KAHAN, 0)
        groups.append((at, {k: at.add(dems[k].z, dems[k].rain, soilT_mode="nonneg") for k in members}))
    assert all(s is not None for _, sl in groups for s in sl.values())
    atl = [g[0] for g in groups]
    steps = 50
    for k in range(steps):
        lemhip.step_group(atl)
        if k % 10 == 9:
            print(f"[config3] step {k + 1}", flush=True)
    sizes = sorted(range(len(dems)), key=lambda k: dems[k].z.size)
    for k in (sizes[0], sizes[len(sizes) // 2]):
        at, sl = next(g for g in groups if k in g[1])
        h, w = dems[k].z.shape
        c = lemhip.LemContext(lemhip.default_params(w, h, 0), lemhip.FP32_KAHAN, 0)
        c.init_state(dems[k].z, dems[k].rain, soilT_mode="nonneg")
        for it in range(steps):
            c.step(-1000 + it)
        for f in ("dem", "fa", "soilT", "stone", "fd"):
            assert np.array_equal(c.download(f), at.download(sl[k], f), equal_nan=True), f"DEM {k}: {f}"
        c.close()
    for at, sl in groups:
        for k, s in sl.items():
            valid = dems[k].z.reshape(-1) != -9999.0
            z, fa = at.download(s, "dem")[valid], at.download(s, "fa")[valid]
            assert np.isfinite(z).all() and (fa >= 0).all(), f"DEM {k}"
        at.close()


def test_serving_alone_equals_in_mix_256_to_4096():
    reqs = serving.make_requests(5, rate=1000.0, seed=23, lo=256, hi=4096, steps=4)
    sides = sorted(r.dem.z.shape[0] for r in reqs)
    print(f"[config4] request sides {sides}", flush=True)
    big = max(r.dem.z.shape[0] for r in reqs)
    srv = serving.Server(big + 1 + 1500, big + 1, lemhip.FP32_KAHAN)
    rep = srv.run(reqs, realtime=False)
    srv.close()
    assert rep["requests"] == len(reqs)
    for r in reqs:
        alone = serving.run_alone(r, lemhip.FP32_KAHAN)
        for f in ("dem", "fa", "soilT"):
            assert np.array_equal(r.result[f], alone[f], equal_nan=True), f"request {r.rid}: {f}"


@pytest.mark.slow
def test_16384_resident_footprint_and_steps():
    side = 16384
    dem = synth.fbm_16384(side=side)
    print("[config5] generated", flush=True)
    free0, total = lemhip.mem_info(0)
    c = lemhip.LemContext(lemhip.default_params(side, side, 0), lemhip.FP32_KAHAN, 0)
    c.init_state(dem.z, None, soilT_mode="nonneg")
    fixed, grow = lemhip.bytes_per_cell(lemhip.FP32_KAHAN, atlas=False)
    c.step(-1000)
    c.step(-999)
    used = free0 - lemhip.mem_info(0)[0]
    print(f"[config5] resident {used / 1e9:.1f} GB of {total / 1e9:.0f} GB; stated {fixed:.0f} + {grow:.0f} B/cell",
          flush=True)
    assert used <= (fixed + grow) * side * side * 1.05 + 2e9
    assert used < 0.5 * total
    z = c.download("dem")
    valid = dem.z.reshape(-1) != -9999.0
    assert np.isfinite(z[valid]).all()
    assert c.summary()["FA_max"] > 0
    c.close()

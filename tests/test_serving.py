"""Serving loop: continuous batching with admission by footprint.  Every
request's result in the mix must equal the same request run alone on its own
context, bit for bit (both precision forms are deterministic)."""
import numpy as np
import pytest

from parallem_amd import lemhip, serving


def test_request_generator_ranges():
    reqs = serving.make_requests(12, rate=50.0, seed=3, lo=64, hi=256, steps=5)
    arr = [r.arrival for r in reqs]
    assert arr == sorted(arr) and arr[0] > 0
    for r in reqs:
        h, w = r.dem.z.shape
        assert 64 <= w <= 256 and h == w
        assert len(r.forcing) == 5
        assert all(300 <= a <= 2500 and -5 <= b <= 25 for a, b in r.forcing)
    again = serving.make_requests(12, rate=50.0, seed=3, lo=64, hi=256, steps=5)
    assert all(np.array_equal(a.dem.z, b.dem.z) and a.forcing == b.forcing for a, b in zip(reqs, again))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [lemhip.FP64, lemhip.FP32_KAHAN])
def test_alone_equals_in_mix_under_continuous_batching(precision):
    reqs = serving.make_requests(6, rate=1000.0, seed=11, lo=48, hi=160, steps=4)
    # room for about two requests at a time: later ones join as earlier ones finish
    srv = serving.Server(256, 256, precision)
    live_counts = []
    orig_step = srv.step

    def counting_step(now_fn):
        live_counts.append(len(srv.live))
        return orig_step(now_fn)
    srv.step = counting_step
    rep = srv.run(reqs, realtime=False)
    srv.close()
    assert rep["requests"] == len(reqs)
    assert max(live_counts) >= 2, live_counts                     # steps overlapped on the device
    admitted_late = [r for r in reqs if r.admitted > min(q.step_done[-1] for q in reqs)]
    assert admitted_late, "no request joined after another finished"
    for r in reqs:
        alone = serving.run_alone(r, precision)
        for f in ("dem", "fa", "soilT"):
            assert np.array_equal(r.result[f], alone[f], equal_nan=True), f"request {r.rid}: {f}"
        for k in ("FA_max", "totE", "totD", "totbio", "outletcellidx", "sinkcount"):
            assert r.result["summary"][k] == alone["summary"][k], f"request {r.rid}: {k}"
    assert rep["p99_step_ms"] >= rep["p50_step_ms"] > 0

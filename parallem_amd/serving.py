"""Serving loop: continuous batching of landscape-evolution solves on one GPU.

A request is one scenario: a DEM plus a climate forcing series (one
(rain mm/yr, temperature C) pair per timestep).  Requests arrive on a Poisson
clock.  Admission is by HBM footprint: the server owns one atlas context whose
grid was sized from an HBM budget, and a request is admitted when a rectangle
for its DEM is free (its footprint is its padded cell count times the
per-cell bytes of the precision form).  Admitted requests keep their state
resident; every server iteration steps all live requests with one launch
sequence, requests join between iterations as soon as room frees, and a
finished request releases its rectangle at once.

Per step each live request gets SetClimateMatrices(k) semantics with its own
forcing (Data.cpp:231-271) before the step and copytolastclimate
(memory_dev.cu:363-373) after it, the canonical rule for forcing series.

Admission walks the arrived requests in arrival order and admits every one
that fits, so a large DEM waiting for room does not hold back small ones
behind it; once the oldest waiting request has been passed over for
`max_bypass` iterations, nothing behind it is admitted until it fits (no
starvation).  The footprint per cell comes from the library
(lemhip_bytes_per_cell): the fixed arrays of an atlas context plus the bound
on the depression-routing arrays.

Latency is the time between a request's consecutive step completions (the
first step counts from admission); throughput is cells stepped per second.
"""
from __future__ import annotations

import dataclasses
import math
import time

import numpy as np

from . import lemhip, synth



def bytes_per_cell(precision: int) -> float:
    """Device bytes per atlas cell of a precision form (state, scratch and
    the bound on the depression-routing arrays), from the library."""
    fixed, grow = lemhip.bytes_per_cell(precision, atlas=True)
    return fixed + grow


@dataclasses.dataclass
class Request:
    rid: int
    dem: synth.Dem
    forcing: list            # [(rain_mm, temp_c)] one per step
    arrival: float = 0.0     # seconds after the start of the run
    # filled by the server
    slot: int | None = None
    admitted: float | None = None
    step_done: list = dataclasses.field(default_factory=list)
    result: dict | None = None

    @property
    def steps(self) -> int:
        return len(self.forcing)

    @property
    def cells(self) -> int:
        return int(self.dem.z.size)


def make_requests(count: int, rate: float, seed: int, lo: int = 256, hi: int = 4096, steps: int = 20) -> list:
    """Poisson arrivals; DEM side log-uniform in [lo, hi]; forcing rain
    U(300, 2500) mm/yr and temperature U(-5, 25) C per step."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    out = []
    for k in range(count):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        side = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        side = max(lo, min(hi, side))
        dem = synth.make_dem(side, seed * 7919 + k, pits=max(1, 40 * side * side // (512 * 512)),
                             plateaus=max(1, 10 * side * side // (512 * 512)), outline=True)
        forcing = [(float(rng.uniform(300, 2500)), float(rng.uniform(-5, 25))) for _ in range(steps)]
        out.append(Request(k, dem, forcing, t))
    return out


class Server:
    def __init__(self, width: int, height: int, precision: int = lemhip.FP32_KAHAN, device: int = 0,
                 hbm_budget: float | None = None, flowtype: int = 0, soilT_mode: str = "nonneg",
                 keep_fields: tuple = ("dem", "fa", "soilT"), max_bypass: int = 8):
        self.precision = precision
        self.bpc = bytes_per_cell(precision)
        self.budget = hbm_budget if hbm_budget is not None else self.bpc * width * height
        self.max_bypass = max_bypass
        self.params = lemhip.default_params(width, height, flowtype)
        self.atlas = lemhip.LemAtlas(self.params, width, height, precision, device)
        self.soilT_mode = soilT_mode
        self.keep = keep_fields
        self.live: dict[int, Request] = {}
        self.used = 0.0
        self.cells_stepped = 0
        self.iterations = 0

    @staticmethod
    def sized_for(max_side: int, precision: int, budget_bytes: float, **kw) -> "Server":
        """An atlas as large as the budget allows, at least one max-size DEM."""
        cells = budget_bytes / bytes_per_cell(precision)
        side = max(max_side + 1, int(math.sqrt(cells)))
        return Server(side, side, precision, hbm_budget=budget_bytes, **kw)

    def footprint(self, r: Request) -> float:
        h, w = r.dem.z.shape
        return self.bpc * (w + 1) * (h + 1)

    def try_admit(self, r: Request, now: float) -> bool:
        if self.used + self.footprint(r) > self.budget:
            return False
        h, w = r.dem.z.shape
        s = self.atlas.open(w, h)
        if s is None:
            return False
        self.atlas.init(s, r.dem.z, r.dem.rain, soilT_mode=self.soilT_mode, iteration=self.params.start_iter)
        r.slot, r.admitted = s, now
        self.live[s] = r
        self.used += self.footprint(r)
        return True

    def step(self, now_fn) -> list:
        """One iteration over every live request; returns the requests that finished."""
        for s, r in self.live.items():
            rain, temp = r.forcing[len(r.step_done)]
            self.atlas.set_climate(s, rain, temp)
        self.atlas.step()
        self.atlas.copy_to_last_climate()
        t = now_fn()
        done = []
        for s, r in list(self.live.items()):
            r.step_done.append(t)
            self.cells_stepped += r.cells
            if len(r.step_done) == r.steps:
                r.result = {f: self.atlas.download(s, f) for f in self.keep}
                r.result["summary"] = self.atlas.summary(s)
                self.atlas.remove(s)
                self.used -= self.footprint(r)
                del self.live[s]
                done.append(r)
        self.iterations += 1
        return done

    def run(self, requests: list, realtime: bool = True) -> dict:
        """Serve until every request is done.  realtime: arrivals follow the
        wall clock; otherwise every request is available at once."""
        pending = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter()
        now = (lambda: time.perf_counter() - t0)
        finished = []
        bypassed = 0   # iterations the oldest waiting request has been passed over
        while pending or self.live:
            t = now()
            arrived = [r for r in pending if not realtime or r.arrival <= t]
            for k, r in enumerate(arrived):
                if self.try_admit(r, t):
                    pending.remove(r)
                    if k == 0:
                        bypassed = 0
                elif k == 0 and bypassed >= self.max_bypass:
                    break      # the oldest has waited long enough: keep the room for it
            if pending and arrived and pending[0] is arrived[0]:
                bypassed += 1
            if not self.live:
                if pending and realtime and pending[0].arrival > t:
                    time.sleep(min(0.01, pending[0].arrival - t))
                    continue
                if pending:
                    raise RuntimeError(f"request {pending[0].rid} ({pending[0].dem.z.shape}) does not fit the atlas")
            finished += self.step(now)
        wall = now()
        return self.report(finished, wall)

    def report(self, finished: list, wall: float) -> dict:
        lat = []
        for r in finished:
            prev = r.admitted
            for t in r.step_done:
                lat.append(t - prev)
                prev = t
        lat = np.array(lat) if lat else np.zeros(1)
        return {"requests": len(finished), "iterations": self.iterations, "wall_s": wall,
                "cells_per_s": self.cells_stepped / wall if wall > 0 else 0.0,
                "p50_step_ms": float(np.percentile(lat, 50) * 1e3), "p99_step_ms": float(np.percentile(lat, 99) * 1e3),
                "mean_live": (sum(r.steps for r in finished) / self.iterations) if self.iterations else 0.0}

    def close(self) -> None:
        self.atlas.close()


def run_alone(r: Request, precision: int, flowtype: int = 0, soilT_mode: str = "nonneg",
              keep_fields: tuple = ("dem", "fa", "soilT")) -> dict:
    """The same request on a single-DEM context, for the alone == in-mix check."""
    h, w = r.dem.z.shape
    ctx = lemhip.LemContext(lemhip.default_params(w, h, flowtype), precision, 0)
    ctx.init_state(r.dem.z, r.dem.rain, soilT_mode=soilT_mode)
    it = ctx.params.start_iter
    for rain, temp in r.forcing:
        ctx.set_climate(rain, temp)
        ctx.step(it)
        ctx.copy_to_last_climate()
        it += 1
    out = {f: ctx.download(f) for f in keep_fields}
    out["summary"] = ctx.summary()
    ctx.close()
    return out


def main():
    import argparse
    import json
    ap = argparse.ArgumentParser(description="serving benchmark: Poisson arrivals, continuous batching")
    ap.add_argument("--rate", type=float, default=10.0, help="requests per second")
    ap.add_argument("--requests", type=int, default=24)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--min-side", type=int, default=256)
    ap.add_argument("--max-side", type=int, default=4096)
    ap.add_argument("--budget-gb", type=float, default=40.0)
    ap.add_argument("--seed", type=int, default=4)
    a = ap.parse_args()
    reqs = make_requests(a.requests, a.rate, a.seed, a.min_side, a.max_side, a.steps)
    srv = Server.sized_for(a.max_side, lemhip.FP32_KAHAN, a.budget_gb * 1e9)
    rep = srv.run(reqs)
    srv.close()
    rep.update({"rate": a.rate, "steps": a.steps, "sides": [a.min_side, a.max_side], "atlas": [srv.atlas.W, srv.atlas.H],
                "precision": "fp32_kahan"})
    print(json.dumps(rep))


if __name__ == "__main__":
    main()

"""Long, large run: one big DEM in the FP32 + compensated form beside the plain
FP64 form on the same GPU, reporting the resident footprint, time per step
and the drift between the two after every block of steps.

Workload (default): one 16384 x 16384 fBm DEM (268 M cells, 5000 pits, 1000
plateaus), 30 m cells, vegetation cover seeded uniform(0, 1): the initial
biomass is cover x carrying capacity (the reference has no cover field, so
this is the mapping used here).  With constant climate the vegetation update
resets biomass to capacity after the first step (updates.cu:128).

    python -m parallem_amd.longrun --side 16384 --steps 20 --check-every 5
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import lemhip, synth


def device_used_gb() -> float:
    free, total = lemhip.mem_info(0)
    return (total - free) / 1e9


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--side", type=int, default=16384)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--check-every", type=int, default=5)
    ap.add_argument("--no-fp64", action="store_true", help="FP32 form only (no drift check)")
    ap.add_argument("--seed-offset", type=int, default=0)
    a = ap.parse_args(argv)
    side = a.side
    t0 = time.perf_counter()
    dem = synth.fbm_16384(seed_offset=a.seed_offset, side=side)
    rng = np.random.default_rng(5)
    cover = rng.uniform(0.0, 1.0, dem.z.shape).reshape(-1)
    print(f"[longrun] generated {side}x{side} in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    base = device_used_gb()
    forms = [("fp32_kahan", lemhip.FP32_KAHAN)] + ([] if a.no_fp64 else [("fp64", lemhip.FP64)])
    ctx, footprint = {}, {}
    for name, prec in forms:
        before = device_used_gb()
        c = lemhip.LemContext(lemhip.default_params(side, side, flowtype=0), prec, 0)
        c.init_state(dem.z, None, soilT_mode="nonneg")
        c.upload("TotB", c.download("TotB") * cover)
        ctx[name] = c
        footprint[name] = device_used_gb() - before
        print(f"[longrun] {name}: {footprint[name]:.1f} GB resident", file=sys.stderr, flush=True)
    mask = ctx["fp32_kahan"].download("mask") == 1
    ms = {name: [] for name in ctx}
    checks = []
    it = -1000
    for k in range(a.steps):
        for name, c in ctx.items():
            if k == 0:
                c.autotune()
            s = time.perf_counter()
            c.step(it)
            ms[name].append((time.perf_counter() - s) * 1e3)
        it += 1
        if "fp64" in ctx and ((k + 1) % a.check_every == 0 or k + 1 == a.steps):
            z32, z64 = ctx["fp32_kahan"].download("dem")[mask], ctx["fp64"].download("dem")[mask]
            d = np.abs(z32 - z64)
            fd32, fd64 = ctx["fp32_kahan"].download("fd")[mask], ctx["fp64"].download("fd")[mask]
            checks.append({"step": k + 1, "max_abs_dz_m": float(d.max()), "p9999_abs_dz_m": float(np.percentile(d, 99.99)),
                           "cells_above_1e-4_m": int(np.sum(d > 1e-4)), "direction_mismatches": int(np.sum(fd32 != fd64))})
            print(f"[longrun] step {k + 1}: {checks[-1]}", file=sys.stderr, flush=True)
    out = {"workload": f"one {side}x{side} fBm DEM (5000 pits, 1000 plateaus at 16384^2), SFD, cover-seeded biomass",
           "cells": side * side, "steps": a.steps, "device_gb_before": base,
           "resident_gb": footprint,
           "ms_per_step": {n: float(np.median(v[1:] if len(v) > 1 else v)) for n, v in ms.items()},
           "cells_per_s": {n: side * side / (float(np.median(v[1:] if len(v) > 1 else v)) * 1e-3) for n, v in ms.items()},
           "tuning": {n: c.tuning() for n, c in ctx.items()},
           "drift": checks}
    for c in ctx.values():
        c.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()

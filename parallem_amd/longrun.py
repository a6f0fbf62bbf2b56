"""Long, large run: one big DEM in the FP32 + compensated form beside the plain
FP64 form on the same GPU, reporting the resident footprint, time per step
and the drift between the two after every block of steps.

Workload (default): one 16384 x 16384 fBm DEM (268 M cells, 5000 pits, 1000
plateaus), 30 m cells, vegetation cover seeded uniform(0, 1): the initial
biomass is cover x carrying capacity (the reference has no cover field, so
this is the mapping used here).  With constant climate the vegetation update
resets biomass to capacity after the first step (updates.cu:128).

    python -m parallem_amd.longrun --side 16384 --steps 1000 --check-every 100

With both forms, directions are compared after every step; each check
reports the drift, and separately the drift off the receiver paths of every
cell whose direction has differed between the forms at some step (a
direction at a near-exact slope tie taken differently sends upstream flow
down another path, so both paths see other discharge from then on).
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


_DIRS = {64: (-1, 0), 128: (-1, 1), 1: (0, 1), 2: (1, 1), 4: (1, 0), 8: (1, -1), 16: (0, -1), 32: (-1, -1)}


def downstream_band(fd, start, W, H):
    """Cells on the single-direction receiver paths from the cells in `start`
    (bool per cell), together with their 8-neighbourhoods (which see the
    changed slopes).  All paths advance together, one link per pass."""
    n = W * H
    idx = np.arange(n, dtype=np.int64)
    rec = np.full(n, -1, np.int64)
    r, c = idx // W, idx % W
    for code, (dy, dx) in _DIRS.items():
        sel = fd == code
        rr, cc = r[sel] + dy, c[sel] + dx
        ok = (rr >= 0) & (rr < H) & (cc >= 0) & (cc < W)
        rec[np.flatnonzero(sel)[ok]] = rr[ok] * W + cc[ok]
    del r, c, idx
    on = np.zeros(n, bool)
    p = np.flatnonzero(start)
    while p.size:
        on[p] = True
        p = rec[p]
        p = p[p >= 0]
        p = np.unique(p[~on[p]])
    g = on.reshape(H, W)
    band = g.copy()
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if dy or dx:
                band[max(dy, 0):H + min(dy, 0), max(dx, 0):W + min(dx, 0)] |= \
                    g[max(-dy, 0):H + min(-dy, 0), max(-dx, 0):W + min(-dx, 0)]
    return band.reshape(-1)


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
    flipped = np.zeros(side * side, bool)   # cells whose direction differed between the forms at some step
    it = -1000
    for k in range(a.steps):
        for name, c in ctx.items():
            if k == 0:
                c.autotune()
            s = time.perf_counter()
            c.step(it)
            ms[name].append((time.perf_counter() - s) * 1e3)
        it += 1
        if "fp64" not in ctx:
            continue
        fd32, fd64 = ctx["fp32_kahan"].download("fd"), ctx["fp64"].download("fd")
        now = mask & (fd32 != fd64)
        flipped |= now
        if (k + 1) % 10 == 0:
            print(f"[longrun] step {k + 1}, {int(now.sum())} direction mismatches", file=sys.stderr, flush=True)
        if (k + 1) % a.check_every == 0 or k + 1 == a.steps:
            z32, z64 = ctx["fp32_kahan"].download("dem"), ctx["fp64"].download("dem")
            d = np.where(mask, np.abs(z32 - z64), 0.0)
            aff = downstream_band(fd64, flipped, side, side)
            off = mask & ~aff
            checks.append({"step": k + 1, "max_abs_dz_m": float(d.max()),
                           "p9999_abs_dz_m": float(np.percentile(d[mask], 99.99)),
                           "cells_above_1e-4_m": int(np.sum(d > 1e-4)),
                           "direction_mismatches": int(now.sum()),
                           "cells_ever_flipped": int(flipped.sum()),
                           "cells_on_flipped_paths": int(aff[mask].sum()),
                           "cells_above_1e-4_m_off_flipped_paths": int(np.sum(d[off] > 1e-4)),
                           "max_abs_dz_off_flipped_paths_m": float(d[off].max())})
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

"""Batched uneven catchments: many DEMs of different sizes stepped together.

Workload (default): 64 fBm catchments with sides log-uniform in [256, 2048]
cells and irregular random-walk outlines (nodata outside), the FP32 +
compensated form, 50 timesteps after one autotuned warmup step.  The DEMs are
shelf-packed into `--groups` atlas grids (largest first), and the atlases are
stepped together with lemhip_step_group (one host thread and one stream per
atlas, so one atlas's latency-bound rounds overlap another's bandwidth-bound
kernels).  Reported: cells/s over the timed steps, p50/p99 step latency, the
packing, and an alone == in-mix check on a sample of the DEMs (each run alone
on a single-DEM context for the same steps, compared bit for bit).

    python -m parallem_amd.batchrun --dems 64 --steps 50 --groups 2 --check 3
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import lemhip, synth


def pack(dems, groups: int, width: int):
    """Assign DEMs (largest first) to the group with the fewest cells, then
    size each group's atlas: shelf rows of the given width."""
    order = sorted(range(len(dems)), key=lambda k: -dems[k].z.size)
    load = [0] * groups
    members = [[] for _ in range(groups)]
    for k in order:
        g = min(range(groups), key=lambda q: load[q])
        members[g].append(k)
        load[g] += dems[k].z.size
    atlases = []
    for g in range(groups):
        # shelf heights follow the packing order; a generous height bound, trimmed by the packer's use
        # (rectangles start at multiples of lemhip.ATLAS_ALIGN)
        a = lemhip.ATLAS_ALIGN
        x, shelf_h, h_total = 0, 0, 0
        for k in members[g]:
            h, w = dems[k].z.shape
            if x + w + 1 > width:
                h_total += -(-shelf_h // a) * a
                x, shelf_h = 0, 0
            x += -(-(w + 1) // a) * a
            shelf_h = max(shelf_h, h + 1)
        h_total += -(-shelf_h // a) * a
        # the library's shelf packer may place DEMs differently from this
        # estimate: one more shelf of the tallest DEM covers the difference
        # (a step covers only the rows its shelves use)
        tallest = max((dems[k].z.shape[0] for k in members[g]), default=0)
        atlases.append((members[g], width, h_total + tallest + 1))
    return atlases


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--dems", type=int, default=64)
    ap.add_argument("--lo", type=int, default=256)
    ap.add_argument("--hi", type=int, default=2048)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--groups", type=int, default=2)
    ap.add_argument("--width", type=int, default=8192)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--check", type=int, default=3, help="DEMs run alone for the alone == in-mix check")
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    a = ap.parse_args(argv)
    prec = lemhip.FP32_KAHAN if a.precision == "fp32" else lemhip.FP64
    t0 = time.perf_counter()
    dems = synth.catchment_batch(a.dems, a.seed, a.lo, a.hi)
    print(f"[batchrun] {len(dems)} DEMs generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    cells = sum(d.z.size for d in dems)
    groups = []
    for members, w, h in pack(dems, a.groups, a.width):
        at = lemhip.LemAtlas(lemhip.default_params(w, h, 0), w, h, prec, 0)
        slots = {}
        for k in members:
            s = at.add(dems[k].z, dems[k].rain, soilT_mode="nonneg")
            if s is None:
                raise RuntimeError(f"DEM {k} {dems[k].z.shape} does not fit a {w}x{h} atlas")
            slots[k] = s
        groups.append((at, slots))
    print(f"[batchrun] packed into {[(g[0].W, g[0].H, len(g[1])) for g in groups]}", file=sys.stderr, flush=True)
    atl = [g[0] for g in groups]
    for at in atl:
        at.autotune()
    lemhip.step_group(atl)           # warmup + tuning
    for at in atl:
        at.profile(True)
    lat = []
    t1 = time.perf_counter()
    for k in range(a.steps):
        s = time.perf_counter()
        lemhip.step_group(atl)
        lat.append(time.perf_counter() - s)
        if (k + 1) % 10 == 0:
            print(f"[batchrun] step {k + 1}/{a.steps}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t1
    prof = [at.profile_read() for at in atl]
    lat = np.array(lat)
    out = {"workload": f"{len(dems)} fBm catchments, sides log-uniform [{a.lo}, {a.hi}], irregular outlines, "
                       f"{a.precision}, {a.steps} steps after one warmup",
           "cells": cells, "valid_cells": int(sum(int((d.z != -9999.0).sum()) for d in dems)),
           "groups": [{"atlas": [at.W, at.H], "dems": len(sl), "tuning": at.tuning()} for at, sl in groups],
           "cells_per_s": cells * a.steps / wall, "ms_per_step": 1e3 * wall / a.steps,
           "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
           "phase_ms_per_step": [{k: round(v[0] / a.steps, 3) for k, v in p.items()} for p in prof]}
    # alone == in-mix on a sample (largest, smallest, median)
    if a.check > 0:
        sizes = sorted(range(len(dems)), key=lambda k: dems[k].z.size)
        pick = sorted({sizes[-1], sizes[0], sizes[len(sizes) // 2]})[:a.check]
        same = {}
        for k in pick:
            at, sl = next(g for g in groups if k in g[1])
            h, w = dems[k].z.shape
            c = lemhip.LemContext(lemhip.default_params(w, h, 0), prec, 0)
            c.init_state(dems[k].z, dems[k].rain, soilT_mode="nonneg")
            for it in range(a.steps + 1):
                c.step(-1000 + it)
            same[int(k)] = all(np.array_equal(c.download(f), at.download(sl[k], f), equal_nan=True)
                               for f in ("dem", "fa", "soilT", "stone", "fd"))
            c.close()
        out["alone_equals_in_mix"] = same
    for at in atl:
        at.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()

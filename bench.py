#!/usr/bin/env python3
"""Benchmark: cells/s per landscape-evolution timestep at 4096^2 on one MI355X.

Workload (the configuration the metric is quoted on): one 4096 x 4096 fBm DEM
(diamond-square, H = 0.8, 0-800 m, 0.5 % tilt, 400 pits, 100 plateaus, 5 %
nodata frame), 30 m cells, per-cell lognormal(0, 0.3) rainfall rescaled to a
1000 mm/yr mean, SFD routing, default parameters.  Soil thickness starts from
the reference formula clamped at 0 (the unclamped formula turns every cell
below 400 m into NaN after one step).

A step is one full FD -> FA -> MOD timestep of the whole DEM with all state
resident in HBM.  `value` is the FP32 + compensated-accumulation form; the
plain FP64 form (bit-exact with the CPU restatement) is timed beside it.
Each step call returns after a stream synchronisation, so the wall clock
around the K timed steps brackets the device work exactly.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cells/s per LEM timestep at 4096² (batched FP32 vs plain); p50/p99 step latency"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per cell-timestep (each persistent field read and written
# once, each hand-off written once and read once), with this build's layouts:
#   FP64 form: dem 8/8, Slopes 8x8/8x8, soilT 8/8, 6 soil/veg fields 48/48,
#              4 climate fields 32/0, fd 1/1 (u8), fa 8/8            -> 306 B
#   FP32 form: dem 8/8, Slopes 8x4/8x4, soilT 8/8, 6 fields 24/24,
#              4 climate fields 16/0, fd 1/1, fa 8/8 (float pair)     -> 178 B
B_ALG = {"fp64": 306, "fp32": 178}

# Algorithmic bytes per cell of the flow-accumulation pass (the phase the
# timing identifies as dominant): donor bits built from fd (1 B), runoff
# weight read (T), fa written once and read once by its receivers (8 + 8 B).
B_FA = {"fp64": 1 + 8 + 16, "fp32": 1 + 4 + 16}
# MOD part 2 kernel (weathering + surface/nutrient/vegetation update), per
# cell: reads fa, ero, inci, geli, depo, temp, rain, last_temp, soilT(8),
# stone, fines, soilM, soilB, nut, TotB, dem(8); writes dz, wC, wP, soilT(8),
# stone, fines, soilM, soilB, nut, TotB, dem(8).
def b_mod2(t):
    fa = 8
    return fa + 7 * t + 8 + 6 * t + 8 + 3 * t + 8 + 6 * t + 8


def gen_dem(side):
    from parallem_amd import synth
    return synth.config_dem(2, side=side)


def run_form(side, dem, precision, steps, warmup, profile):
    from parallem_amd import lemhip
    ctx = lemhip.LemContext(lemhip.default_params(side, side, flowtype=0), precision, 0)
    ctx.init_state(dem.z, dem.rain, soilT_mode="nonneg")
    it = -1000
    for _ in range(warmup):
        ctx.step(it)
        it += 1
    if profile:
        ctx.profile(True)
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        ctx.step(it)
        lat.append(time.perf_counter() - a)
        it += 1
    total = time.perf_counter() - t0
    prof = ctx.profile_read() if profile else None
    summ = ctx.summary()
    ctx.close()
    return total, np.array(lat), prof, summ


def cpu_baseline(side, threads):
    from oracle import oracle as O
    from parallem_amd import synth
    dem = synth.config_dem(2, side=side)
    O.lib(True).orc_set_threads(threads)
    o = O.Oracle(dem.z, dem.rain, O.default_params(side, side), openmp=True, soilT_mode="nonneg")
    t0 = time.perf_counter()
    o.step(-1000)
    dt = time.perf_counter() - t0
    return side * side / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--side", type=int, default=4096)
    ap.add_argument("--no-plain", action="store_true", help="skip the FP64 form")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-side", type=int, default=2048)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from parallem_amd import lemhip
    side = args.side
    N = side * side
    print(f"[bench] generating {side}x{side} DEM", file=sys.stderr, flush=True)
    dem = gen_dem(side)

    print("[bench] FP32+Kahan form", file=sys.stderr, flush=True)
    t32, lat32, prof32, sum32 = run_form(side, dem, lemhip.FP32_KAHAN, args.steps, args.warmup, True)
    value = N * args.steps / t32
    plain = None
    if not args.no_plain:
        print("[bench] plain FP64 form", file=sys.stderr, flush=True)
        t64, lat64, prof64, _ = run_form(side, dem, lemhip.FP64, args.steps, args.warmup, True)
        plain = {"value": N * args.steps / t64, "ms_per_step": 1e3 * t64 / args.steps,
                 "p50_ms": float(np.percentile(lat64, 50) * 1e3), "p99_ms": float(np.percentile(lat64, 99) * 1e3),
                 "phase_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in prof64.items()},
                 "step_roofline_frac": N * args.steps / t64 * B_ALG["fp64"] / (HBM_PEAK_GBS * 1e9)}

    # dominant phase of the FP32 run by live HIP-event time
    phases = {k: v for k, v in prof32.items() if k not in ("step", "fd")}
    dom = max(phases, key=lambda k: phases[k][0])
    dom_ms = phases[dom][0]
    if dom == "flow_accum":
        bytes_per_cell = B_FA["fp32"]
    elif dom == "mod2":
        bytes_per_cell = b_mod2(4)
    else:
        bytes_per_cell = None
    roof = {"kernel": dom, "bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS, "traffic": None}
    if bytes_per_cell is not None and dom_ms > 0:
        ach = bytes_per_cell * N * args.steps / (dom_ms * 1e-3) / 1e9
        roof.update({"achieved": ach, "frac": ach / HBM_PEAK_GBS,
                     "bytes_per_cell": bytes_per_cell, "ms_per_step": dom_ms / args.steps,
                     "launches_per_step": phases[dom][1] / args.steps})
    else:
        roof.update({"achieved": None, "frac": None, "ms_per_step": dom_ms / args.steps,
                     "launches_per_step": phases[dom][1] / args.steps,
                     "note": "iterative phase without a single-pass byte count"})

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        print(f"[bench] CPU baseline (oracle, OpenMP {threads} threads, {args.cpu_side}^2, 1 step)",
              file=sys.stderr, flush=True)
        v, dt = cpu_baseline(args.cpu_side, threads)
        cpu = {"value": v, "unit": "cells/s", "cores": threads, "kind": "port",
               "sample": f"one FP64 timestep of a {args.cpu_side}x{args.cpu_side} DEM from the same generator "
                         f"(C restatement, OpenMP), {dt:.1f} s", "host": platform.processor() or platform.machine()}

    info = lemhip.device_info(0)
    out = {
        "metric": METRIC, "value": value, "unit": "cells/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * t32 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 dem/soilT, f32x2 fa)",
        "data": "synthetic",
        "config": {"workload": f"single {side}x{side} fBm DEM (400 pits, 100 plateaus, lognormal rain), "
                               "SFD, FP32+Kahan timestep", "side": side, "cells": N,
                   "flowdirectiontype": 0, "precision": "fp32_kahan"},
        "p50_ms": float(np.percentile(lat32, 50) * 1e3), "p99_ms": float(np.percentile(lat32, 99) * 1e3),
        "step_roofline_frac": value * B_ALG["fp32"] / (HBM_PEAK_GBS * 1e9),
        "phase_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in prof32.items()},
        "roofline": roof,
        "plain_fp64": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "cpu_baseline": cpu,
        "device": info,
        "last_step": {k: sum32[k] for k in ("sinkcount", "plateau_rounds1", "plateau_rounds2",
                                             "fa_uncomputed", "sed_uncomputed")},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

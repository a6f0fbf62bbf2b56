#!/usr/bin/env python3
"""Benchmark: cells/s per landscape-evolution timestep on 4096^2 DEMs, one MI355X.

Workloads (all with state resident in HBM, 30 m cells, SFD routing, default
parameters, soil thickness from the reference formula clamped at 0 — the
unclamped formula turns every cell below 400 m into NaN after one step):

* value — batched FP32 + compensated accumulation: B (default 8) independent
  4096 x 4096 fBm DEMs (diamond-square H = 0.8, 0-800 m, 0.5 % tilt, 400 pits,
  100 plateaus, 5 % nodata frame, per-cell lognormal(0, 0.3) rain rescaled to
  a 1000 mm/yr mean; seeds differ per DEM) split over --groups (default 3)
  atlas grids, stepped together on one stream each;
* single_fp32 — one such DEM alone, same precision form;
* plain_fp64 — one such DEM alone in the plain FP64 form (bit-exact with the
  CPU restatement): the reference's own arithmetic.

Before timing, one warmup step runs the launch autotuner (tile shape and
launch batch of the dependency engine, plateau rounds per launch), which
never changes results.  A step is one full FD -> FA -> MOD timestep of every
DEM.  Every step call
returns after a stream synchronisation, so the wall clock around the K timed
steps brackets the device work exactly.  cells counted = sum of W*H over DEMs.
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
SIDE = 4096

# Algorithmic bytes per cell-timestep of the whole step (each persistent field
# read and written once, each hand-off written once and read once), with the
# layouts used here (u8 fd instead of the reference's i32):
#   FP64 form: dem 8/8, Slopes 8x8/8x8, soilT 8/8, 6 soil/veg fields 48/48,
#              4 climate fields 32/0, fd 1/1, fa 8/8                   -> 306 B
#   FP32 form: dem 8/8, Slopes 8x4/8x4, soilT 8/8, 6 fields 24/24,
#              4 climate fields 16/0, fd 1/1, fa 8/8 (float pair)       -> 178 B
B_ALG = {"fp64": 306, "fp32": 178}


def kernel_bytes(kernel, t):
    """Algorithmic HBM bytes per cell for one pass of a kernel family over the
    grid (t = bytes of a bulk FP element: 8 FP64 form, 4 FP32 form).

    fa_rounds  (k_dag_round, flow accumulation): donor bits 1 + fd 1 + runoff
               weight t read, fa 8 written (f64, or float pair).
    sed_rounds (k_dag_round, sediment routing): donor bits 1 + fd 1, ero, inci,
               geli, stone t each + soilT 8 read; depo, stone, fines t each +
               soilT 8 written.
    mod2       (k_mod2): fa 8, ero, inci, geli, depo, temp, rain, last_temp,
               stone, fines, soilM, soilB, nut, TotB and the two calcite
               terms t each, soilT 8, dem 8 read; dz, wC, wP, stone, fines,
               soilM, soilB, nut, TotB t each, soilT 8, dem 8 written.
    mod1       (k_mod1): Slopes 8t, fa 8, fines, nut, stone, TotB, temp t each,
               soilT 8 read; SlopePtr, ero, inci, geli, depo t each written.
    flowdirs   (k_flowdirs): dem 8 (stencil, read once from HBM), mask 1 read;
               fd, SFD 1 each + downhill Slopes entries (at most 8t) written.
    plat_rounds (k_plat_block, plateau routing on the raised surface): surface
               8, mask 1, equal-height bits 1 read; fd, SFD 1 each and the
               distance 4 written, once per cell.
    spill_relax (the Borůvka phases of the spill heights, k_bv_*) is counted
               per basin graph, not per cell: see graph_bytes.
    In the FP32 form sed_rounds is the engine's first round over every tile
    (one pass, the bytes above); sed_contract (k_cs_round, the contraction of
    the cells that round leaves open) is counted per open cell: parameters
    2t, status word 4 and donor bits 1 read, stone t and status word 4
    written.
    """
    return {
        "fa_rounds": 2 + t + 8,
        "sed_rounds": 2 + 4 * t + 8 + 3 * t + 8,
        "mod2": 8 + 15 * t + 8 + 8 + 9 * t + 8 + 8,
        "mod1": 8 * t + 8 + 5 * t + 8 + 5 * t,
        "flowdirs": 8 + 1 + 2 + 8 * t,
        "plat_rounds": 8 + 1 + 1 + 2 + 4,
        "sed_contract": 2 * t + 4 + 1 + t + 4,
    }.get(kernel)


def graph_bytes(nedges, nlabels):
    """Algorithmic bytes of the spill-height pass: the basin-pair edge list
    (two labels 4 each, weight 8) read once, each label's spill height 8
    written once."""
    return 16 * nedges + 8 * nlabels


# kernel-level timing slots (HIP events around exactly these launches)
KERNEL_SLOTS = ("fa_rounds", "sed_rounds", "mod1", "mod2", "flowdirs", "plat_rounds", "spill_relax", "sed_contract")
# rocprofv3 kernel names of those slots (the round kernels carry the tuned tile shape)
def rocprof_name(slot, tuning, t=4):
    tw, th = tuning["dag_tile"]
    ts = "float" if t == 4 else "double"
    sw, sh = (32, 32) if t == 4 and tuning.get("fp32_tree_sediment") else (tw, th)   # fixed first-round shape
    return {"fa_rounds": f"k_dag_round<{tw}, {th}, FaCell", "sed_rounds": f"k_dag_round<{sw}, {sh}, SedCell",
            "sed_contract": "k_cs_round",
            "mod1": "k_mod1", "mod2": "k_mod2", "flowdirs": "k_flowdirs",
            # pass 2 (raised surface, no slope marks) holds nearly all the rounds
            "plat_rounds": f"k_plat_block<{tuning['plateau_rounds_per_launch']}, {ts}, false>",
            "spill_relax": "k_bv_min_w"}[slot]


def dems(count):
    from parallem_amd import synth
    return [synth.fbm_4096(seed_offset=k, side=SIDE) for k in range(count)]


def timed(step, steps, warmup, handle, autotune=True):
    if autotune:
        # the first warmup step times the launch candidates on this data and keeps the fastest
        handle.autotune()
    for _ in range(max(warmup, 1 if autotune else 0)):
        step()
    handle.profile(True)
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        step()
        lat.append(time.perf_counter() - a)
    total = time.perf_counter() - t0
    return total, np.array(lat), handle.profile_read()


def run_single(dem, precision, steps, warmup, tuning=None):
    from parallem_amd import lemhip
    ctx = lemhip.LemContext(lemhip.default_params(SIDE, SIDE, flowtype=0), precision, 0)
    ctx.init_state(dem.z, dem.rain, soilT_mode="nonneg")
    if tuning:
        ctx.set_tuning(**tuning)
    it = [-1000]

    def step():
        ctx.step(it[0])
        it[0] += 1
    total, lat, prof = timed(step, steps, warmup, ctx, autotune=tuning is None)
    summ = ctx.summary()
    summ["tuning"] = ctx.tuning()
    ctx.close()
    return total, lat, prof, summ


class Group:
    """Several atlases stepped together (lemhip_step_group), driven through
    the same timing helper as one context."""

    def __init__(self, atlases):
        self.atlases = atlases

    def autotune(self):
        for a in self.atlases:
            a.autotune()

    def profile(self, on):
        for a in self.atlases:
            a.profile(on)

    def profile_read(self):
        """Phase times summed over the atlases (each on its own stream)."""
        out = {}
        for a in self.atlases:
            for k, (ms, n) in a.profile_read().items():
                m0, n0 = out.get(k, (0.0, 0))
                out[k] = (m0 + ms, n0 + n)
        return out

    def step(self):
        from parallem_amd import lemhip
        lemhip.step_group(self.atlases)


def run_batched(ds, precision, steps, warmup, tuning=None, groups=1):
    """The DEMs split over `groups` atlases (each a grid of DEM rectangles with
    the fewest empty slots, then the squarest), stepped together."""
    from parallem_amd import lemhip
    parts = [ds[g::groups] for g in range(groups)]
    atlases, slots, dims = [], [], []
    for part in parts:
        n = len(part)
        cols = min(range(1, n + 1), key=lambda c: (c * -(-n // c), abs(c - -(-n // c))))
        rows = -(-n // cols)
        aw, ah = lemhip.atlas_span([SIDE] * cols), lemhip.atlas_span([SIDE] * rows)
        atlas = lemhip.LemAtlas(lemhip.default_params(aw, ah, flowtype=0), aw, ah, precision, 0)
        sl = [atlas.add(d.z, d.rain, soilT_mode="nonneg") for d in part]
        assert all(s is not None for s in sl)
        if tuning:
            atlas.set_tuning(**tuning)
        atlases.append(atlas)
        slots.append(sl)
        dims.append((aw, ah))
    grp = Group(atlases)
    total, lat, prof = timed(grp.step, steps, warmup, grp, autotune=tuning is None)
    summ = [a.summary(s) for a, sl in zip(atlases, slots) for s in sl]
    summ[0]["tuning"] = atlases[0].tuning()
    # basin graphs of the last step (edge and label counts are per atlas)
    firsts = [a.summary(sl[0]) for a, sl in zip(atlases, slots)]
    graph = (sum(f["nedges"] for f in firsts), sum(f["nlabels"] for f in firsts),
             sum(f["sed_open"] for f in firsts))
    for a in atlases:
        a.close()
    return total, lat, prof, summ, dims, graph


def phase_ms(prof, steps):
    return {k: round(v[0] / steps, 3) for k, v in prof.items()}


def roofline(prof, steps, cells, t, tuning, graph):
    """Dominant kernel by live HIP-event time; achieved = algorithmic bytes
    per launch / average launch duration (= bytes of all its launches / their
    total time).  graph: (edges, labels) of the basin graphs of one step and
    the sediment cells the contraction took over."""
    dom = max(KERNEL_SLOTS, key=lambda k: prof[k][0])
    ms, launches = prof[dom]
    if dom == "spill_relax":
        total_bytes = graph_bytes(*graph[:2]) * steps
        b = graph_bytes(*graph[:2]) / cells
    elif dom == "sed_contract":
        total_bytes = kernel_bytes(dom, t) * graph[2] * steps
        b = total_bytes / steps / cells
    else:
        b = kernel_bytes(dom, t)
        # the round kernels share one pass over the grid per step; the others
        # make one pass per launch
        passes = steps if dom in ("fa_rounds", "sed_rounds", "plat_rounds") else launches
        total_bytes = b * cells * passes
    ach = total_bytes / (ms * 1e-3) / 1e9 if ms > 0 else None
    return {"kernel": rocprof_name(dom, tuning, t), "bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
            "achieved": ach, "frac": (ach / HBM_PEAK_GBS) if ach else None, "traffic": None,
            "bytes_per_cell": b, "launches_per_step": launches / steps,
            "avg_launch_us": 1e3 * ms / launches if launches else None,
            "bytes_per_launch": total_bytes / launches if launches else None,
            "ms_per_step": ms / steps}


def traffic_file(batch):
    return os.path.join(ROOT, "profiles", f"batched_{batch}x4096_fp32_traffic.json")


def measured_traffic(kernel, tuning, batch):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE /
    WRITE_SIZE counter runs of the same workload and tuning (calibrated
    against a copy kernel of known size), or None."""
    try:
        with open(traffic_file(batch)) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    # the kernel's name carries the launch shape that matters (tile shape,
    # rounds per launch), so a name match is a match of that shape
    for name, v in t.get("kernels", {}).items():
        if kernel in name:
            return v["fetch_bytes_per_launch"] + v["write_bytes_per_launch"]
    return None


def cpu_step(side, threads):
    """One FP64 timestep of the C restatement on a side^2 DEM from the 4096^2
    generator: cells/s and seconds.  threads 0: the serial build."""
    from oracle import oracle as O
    from parallem_amd import synth
    dem = synth.fbm_4096(side=side)
    if threads:
        O.lib(True).orc_set_threads(threads)
    o = O.Oracle(dem.z, dem.rain, O.default_params(side, side), openmp=threads > 0, soilT_mode="nonneg")
    t0 = time.perf_counter()
    o.step(-1000)
    dt = time.perf_counter() - t0
    return side * side / dt, dt


def cpu_baseline(side, serial_side, threads):
    """OpenMP build on the bench's DEM size, serial build on a smaller sample."""
    v, dt = cpu_step(side, threads)
    vs, dts = cpu_step(serial_side, 0)
    return {"value": v, "unit": "cells/s", "cores": threads, "kind": "port",
            "sample": f"one FP64 timestep of a {side}x{side} DEM from the same generator (C restatement, "
                      f"OpenMP, {threads} threads), {dt:.1f} s",
            "parallel_parts": "FlowDirs, plateau rounds, runoff and the elementwise erosion/weathering/update "
                              "loops; labels, spill heights (Kruskal), flow accumulation and sediment routing "
                              "run serially",
            "serial": {"value": vs, "unit": "cells/s", "cores": 1,
                       "sample": f"one FP64 timestep of a {serial_side}x{serial_side} DEM, serial build, {dts:.1f} s"},
            "host": platform.processor() or platform.machine()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=8, help="4096^2 DEMs in the batched form")
    ap.add_argument("--groups", type=int, default=3,
                    help="atlases the batch is split over, stepped together on their own streams")
    ap.add_argument("--no-plain", action="store_true", help="skip the single-DEM forms")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-side", type=int, default=4096)
    ap.add_argument("--cpu-serial-side", type=int, default=2048)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB three times (a known-size kernel for counter calibration)")
    ap.add_argument("--tuning", default=None,
                    help="tw,th,batch,k: fixed launch tuning instead of the autotuner (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "runs on exactly one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from parallem_amd import lemhip
    if args.copy_calibration > 0:
        import torch
        a = torch.ones(args.copy_calibration * 262144, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    fixed = None
    if args.tuning:
        tw, th, bt, k = (int(x) for x in args.tuning.split(","))
        fixed = dict(dag_tile=(tw, th), dag_batch=bt, plateau_rounds_per_launch=k)
    print(f"[bench] generating {args.batch} DEMs of {SIDE}x{SIDE}", file=sys.stderr, flush=True)
    ds = dems(max(args.batch, 1))
    cells_b = args.batch * SIDE * SIDE

    print("[bench] batched FP32+Kahan form", file=sys.stderr, flush=True)
    tb, latb, profb, summb, atlas_dims, graphb = run_batched(ds[:args.batch], lemhip.FP32_KAHAN, args.steps,
                                                             args.warmup, tuning=fixed, groups=args.groups)
    value = cells_b * args.steps / tb

    single32 = plain = None
    if not args.no_plain:
        print("[bench] single-DEM FP32+Kahan form", file=sys.stderr, flush=True)
        t32, lat32, prof32, s32 = run_single(ds[0], lemhip.FP32_KAHAN, args.steps, args.warmup, tuning=fixed)
        single32 = {"value": SIDE * SIDE * args.steps / t32, "ms_per_step": 1e3 * t32 / args.steps,
                    "p50_ms": float(np.percentile(lat32, 50) * 1e3), "p99_ms": float(np.percentile(lat32, 99) * 1e3),
                    "phase_ms_per_step": phase_ms(prof32, args.steps), "tuning": s32["tuning"]}
        print("[bench] single-DEM plain FP64 form", file=sys.stderr, flush=True)
        t64, lat64, prof64, s64 = run_single(ds[0], lemhip.FP64, args.steps, args.warmup, tuning=fixed)
        plain = {"value": SIDE * SIDE * args.steps / t64, "ms_per_step": 1e3 * t64 / args.steps,
                 "p50_ms": float(np.percentile(lat64, 50) * 1e3), "p99_ms": float(np.percentile(lat64, 99) * 1e3),
                 "phase_ms_per_step": phase_ms(prof64, args.steps),
                 "step_roofline_frac": SIDE * SIDE * args.steps / t64 * B_ALG["fp64"] / (HBM_PEAK_GBS * 1e9),
                 "roofline": roofline(prof64, args.steps, SIDE * SIDE, 8, s64["tuning"],
                                      (s64["nedges"], s64["nlabels"], s64["sed_open"])), "tuning": s64["tuning"],
                 "last_step": {k: s64[k] for k in ("sinkcount", "plateau_rounds1", "plateau_rounds2",
                                                    "fa_uncomputed", "sed_uncomputed")}}

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        print(f"[bench] CPU baseline (oracle: OpenMP {threads} threads on {args.cpu_side}^2, serial on "
              f"{args.cpu_serial_side}^2, 1 step each)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_side, args.cpu_serial_side, threads)

    roof = roofline(profb, args.steps, sum(w * h for w, h in atlas_dims), 4, summb[0]["tuning"], graphb)
    roof["traffic"] = measured_traffic(roof["kernel"], summb[0]["tuning"], args.batch)
    if roof["traffic"] is not None:
        roof["traffic_note"] = "HBM bytes per launch, rocprofv3 FETCH_SIZE + WRITE_SIZE (profiles/)"
    out = {
        "metric": METRIC, "value": value, "unit": "cells/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * tb / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 dem/soilT, f32x2 fa)",
        "data": "synthetic",
        "config": {"workload": f"batch of {args.batch} independent {SIDE}x{SIDE} fBm DEMs (400 pits, 100 plateaus, "
                               f"lognormal rain each) in {args.groups} atlas grid(s) stepped together, SFD, "
                               "FP32+Kahan timestep",
                   "dems": args.batch, "side": SIDE, "cells": cells_b, "atlases": [list(d) for d in atlas_dims],
                   "flowdirectiontype": 0, "precision": "fp32_kahan"},
        "p50_ms": float(np.percentile(latb, 50) * 1e3), "p99_ms": float(np.percentile(latb, 99) * 1e3),
        "step_roofline_frac": value * B_ALG["fp32"] / (HBM_PEAK_GBS * 1e9),
        "phase_ms_per_step": phase_ms(profb, args.steps),
        "roofline": roof,
        "single_fp32": single32,
        "plain_fp64": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "cpu_baseline": cpu,
        "device": lemhip.device_info(0),
        "tuning": summb[0]["tuning"],
        "last_step": [{k: s[k] for k in ("sinkcount", "plateau_rounds1", "plateau_rounds2", "fa_uncomputed",
                                         "sed_uncomputed", "totwater")} for s in summb[:1]],
        # input diagnostics per DEM: valid cells the reference's label kernels
        # also reach from threads past a row end (0: every schedule gives these
        # labels), and valid cells whose formula soil thickness is negative
        # (clamped to 0 in this workload)
        "inputs": {"phantom_cells": [s["phantom_cells"] for s in summb],
                   "soilT_negative_fraction": [round(s["soilT_negative"] / int((d.z != -9999.0).sum()), 4)
                                               for s, d in zip(summb, ds)]},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

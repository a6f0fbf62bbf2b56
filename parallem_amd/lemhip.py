"""Python front end of liblemhip.so (the C-ABI in include/lemhip.h).

Mirrors the reference driver's per-step calls (lem.cpp:83-97) on a context
that keeps every field resident on the GPU:

    cuFlowDirection(ctx, i)      routeflow.cu:686   -> lemhip_flow_direction
    computeRunOffWeights(ctx)    runoffweight.cu:61 -> lemhip_runoff
    correctmfdflow(ctx, i)       mfd_simple.cu:238  -> lemhip_flow_accum
    erosionGPU(ctx, i)           erosion.cu:69      -> lemhip_erosion

There is no CPU fallback: importing works without a GPU (so the library can
be inspected), but creating a context needs the HIP library and a device, and
any failure raises LemError.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblemhip.so")

FP64 = 0
FP32_KAHAN = 1

FIELDS = {
    "dem": 0, "mask": 1, "Slopes": 2, "SlopePtr": 3, "soilT": 4, "soilM": 5, "stone": 6, "fines": 7,
    "nut": 8, "soilB": 9, "TotB": 10, "rain": 11, "temp": 12, "last_rain": 13, "last_temp": 14,
    "fd": 15, "SFD": 16, "runoff": 17, "fa": 18, "ero": 19, "inci": 20, "geli": 21, "depo": 22,
    "dz": 23, "wC": 24, "wP": 25,
}
INT_FIELDS = {"mask", "fd", "SFD"}

EXPORTS = [
    "lemhip_create", "lemhip_destroy", "lemhip_last_error", "lemhip_init_state", "lemhip_upload",
    "lemhip_download", "lemhip_get_outlet", "lemhip_set_outlet", "lemhip_flow_direction", "lemhip_runoff",
    "lemhip_flow_accum", "lemhip_erosion", "lemhip_step", "lemhip_run", "lemhip_set_climate",
    "lemhip_copy_to_last_climate", "lemhip_summary", "lemhip_math_eval", "lemhip_device_info",
    "lemhip_profile_enable", "lemhip_profile_read", "lemhip_create_atlas", "lemhip_slot_open",
    "lemhip_slot_init", "lemhip_slot_close", "lemhip_slot_rect", "lemhip_slot_upload", "lemhip_slot_download",
    "lemhip_slot_get_outlet", "lemhip_slot_set_outlet", "lemhip_slot_set_climate", "lemhip_slot_summary",
    "lemhip_step_slots", "lemhip_get_tuning", "lemhip_set_tuning", "lemhip_autotune",
    "lemhip_mem_info", "lemhip_bytes_per_cell", "lemhip_snapshot_size", "lemhip_snapshot", "lemhip_restore",
    "lemhip_step_group",
]

PROF_SLOTS = ["fd", "flowdirs", "plateau", "depression", "runoff", "flow_accum", "mod1", "sediment",
              "mod2", "reductions", "step", "fa_rounds", "sed_rounds", "plat_rounds", "spill_relax",
              "sed_contract"]


class LemTuning(C.Structure):
    _fields_ = [("dag_tile_w", C.c_int), ("dag_tile_h", C.c_int), ("dag_batch", C.c_int),
                ("plateau_rounds_per_launch", C.c_int), ("tuned_ms", C.c_double * 16),
                ("fp32_integer_fa", C.c_int), ("use_graphs", C.c_int), ("fp32_tree_sediment", C.c_int)]

    def as_dict(self) -> dict:
        return {"dag_tile": [self.dag_tile_w, self.dag_tile_h], "dag_batch": self.dag_batch,
                "plateau_rounds_per_launch": self.plateau_rounds_per_launch,
                "tuned_ms": [round(x, 3) for x in self.tuned_ms], "fp32_integer_fa": self.fp32_integer_fa,
                "use_graphs": self.use_graphs, "fp32_tree_sediment": self.fp32_tree_sediment}


class LemError(RuntimeError):
    pass


class LemParams(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("cellsize", C.c_double), ("flowdirectiontype", C.c_int),
        ("stoneval", C.c_double), ("fineval", C.c_double), ("soilTval", C.c_double),
        ("soilMval", C.c_double), ("soilBval", C.c_double), ("nutval", C.c_double),
        ("soil_struct", C.c_int), ("profile_permeability", C.c_int),
        ("Qthreshold", C.c_double), ("Gslope", C.c_double), ("kappa", C.c_double),
        ("Gflowlim", C.c_double), ("Flow_A", C.c_double), ("Flow_B", C.c_double),
        ("max_incision", C.c_double), ("ddk", C.c_double), ("dck", C.c_double), ("dgk", C.c_double),
        ("temp_inflex", C.c_double), ("temp_sens", C.c_double),
        ("rain", C.c_double), ("temp", C.c_double),
        ("runiniter", C.c_int), ("start_iter", C.c_int),
    ]


class LemSummary(C.Structure):
    _fields_ = [
        ("FA_max", C.c_double), ("totE", C.c_double), ("totI", C.c_double), ("totG", C.c_double),
        ("totD", C.c_double), ("totweatherC", C.c_double), ("totweatherP", C.c_double),
        ("totbio", C.c_double),
        ("outletcellidx", C.c_int), ("sinkcount", C.c_int), ("sinks_pass1", C.c_int),
        ("nlabels", C.c_int), ("nedges", C.c_int), ("plateau_rounds1", C.c_int),
        ("plateau_rounds2", C.c_int), ("fa_uncomputed", C.c_int), ("sed_uncomputed", C.c_int),
        ("totwater", C.c_double), ("phantom_cells", C.c_int), ("soilT_negative", C.c_int),
        ("sed_open", C.c_int),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


def default_params(width: int, height: int, flowtype: int = 0, cellsize: float = 30.0,
                   runiniter: int = 10) -> LemParams:
    """default.par:3-35 on a width x height grid of `cellsize` m cells."""
    return LemParams(width=width, height=height, cellsize=cellsize, flowdirectiontype=flowtype,
                     stoneval=20.0, fineval=80.0, soilTval=1.0, soilMval=0.5, soilBval=0.0, nutval=10.0,
                     soil_struct=2, profile_permeability=3, Qthreshold=6.0, Gslope=1.0, kappa=0.00001,
                     Gflowlim=600.0, Flow_A=0.00001, Flow_B=1.3, max_incision=0.001, ddk=0.9, dck=0.4,
                     dgk=0.9, temp_inflex=6.0, temp_sens=6.0, rain=600.0, temp=10.0,
                     runiniter=runiniter, start_iter=-1000)


_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    """Load liblemhip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LemError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                       "(make -C parallem_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.lemhip_create.argtypes = [C.POINTER(LemParams), C.c_int, C.c_int, C.POINTER(vp)]
    L.lemhip_destroy.argtypes = [vp]
    L.lemhip_last_error.argtypes = [vp]
    L.lemhip_last_error.restype = C.c_char_p
    L.lemhip_init_state.argtypes = [vp, vp, vp, C.c_int]
    L.lemhip_upload.argtypes = [vp, C.c_int, vp]
    L.lemhip_download.argtypes = [vp, C.c_int, vp]
    L.lemhip_get_outlet.argtypes = [vp, C.POINTER(C.c_int)]
    L.lemhip_set_outlet.argtypes = [vp, C.c_int]
    for f in ("lemhip_flow_direction", "lemhip_flow_accum", "lemhip_erosion", "lemhip_step"):
        getattr(L, f).argtypes = [vp, C.c_int]
    L.lemhip_runoff.argtypes = [vp]
    L.lemhip_run.argtypes = [vp, C.c_int, C.c_int]
    L.lemhip_set_climate.argtypes = [vp, C.c_double, C.c_double]
    L.lemhip_copy_to_last_climate.argtypes = [vp]
    L.lemhip_summary.argtypes = [vp, C.POINTER(LemSummary)]
    L.lemhip_math_eval.argtypes = [C.c_int, C.c_int, vp, vp, vp]
    L.lemhip_device_info.argtypes = [C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.lemhip_profile_enable.argtypes = [vp, C.c_int]
    L.lemhip_profile_read.argtypes = [vp, C.c_int, vp, vp]
    ip = C.POINTER(C.c_int)
    L.lemhip_create_atlas.argtypes = [C.POINTER(LemParams), C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]
    L.lemhip_slot_open.argtypes = [vp, C.c_int, C.c_int, ip]
    L.lemhip_slot_init.argtypes = [vp, C.c_int, vp, vp, C.c_int]
    L.lemhip_slot_close.argtypes = [vp, C.c_int]
    L.lemhip_slot_rect.argtypes = [vp, C.c_int, ip, ip, ip, ip]
    L.lemhip_slot_upload.argtypes = [vp, C.c_int, C.c_int, vp]
    L.lemhip_slot_download.argtypes = [vp, C.c_int, C.c_int, vp]
    L.lemhip_slot_get_outlet.argtypes = [vp, C.c_int, ip]
    L.lemhip_slot_set_outlet.argtypes = [vp, C.c_int, C.c_int]
    L.lemhip_slot_set_climate.argtypes = [vp, C.c_int, C.c_double, C.c_double]
    L.lemhip_slot_summary.argtypes = [vp, C.c_int, C.POINTER(LemSummary)]
    L.lemhip_step_slots.argtypes = [vp, vp]
    L.lemhip_get_tuning.argtypes = [vp, C.POINTER(LemTuning)]
    L.lemhip_set_tuning.argtypes = [vp, C.POINTER(LemTuning)]
    L.lemhip_autotune.argtypes = [vp]
    dp = C.POINTER(C.c_double)
    L.lemhip_mem_info.argtypes = [C.c_int, dp, dp]
    L.lemhip_bytes_per_cell.argtypes = [C.c_int, C.c_int, dp, dp]
    L.lemhip_snapshot_size.argtypes = [vp, C.c_int, C.POINTER(C.c_size_t)]
    L.lemhip_snapshot.argtypes = [vp, C.c_int, vp, C.c_size_t]
    L.lemhip_restore.argtypes = [vp, C.c_int, vp, C.c_size_t]
    L.lemhip_step_group.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(C.c_int), C.POINTER(C.POINTER(C.c_int))]
    _lib = L
    return L


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _buf(name: str, n: int) -> np.ndarray:
    if name == "Slopes":
        return np.empty(8 * n, np.float64)
    return np.empty(n, np.int32 if name in INT_FIELDS else np.float64)


class _Handle:
    def _open(self, create, what: str) -> None:
        h = C.c_void_p()
        rc = create(C.byref(h))
        self.h = h
        if rc != 0:
            msg = self.L.lemhip_last_error(h).decode() if h.value else f"{what} failed"
            if h.value:
                self.L.lemhip_destroy(h)
            self.h = C.c_void_p()
            raise LemError(f"{what}: {rc} {msg}")

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            raise LemError(f"{what}: {rc} {self.L.lemhip_last_error(self.h).decode()}")

    def close(self) -> None:
        if getattr(self, "h", None) and self.h.value:
            self.L.lemhip_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def profile(self, on: bool) -> None:
        """Start (and reset) or stop live HIP-event timing of the phases."""
        self._check(self.L.lemhip_profile_enable(self.h, 1 if on else 0), "profile_enable")

    def profile_read(self) -> dict:
        n = len(PROF_SLOTS)
        ms = np.zeros(n, np.float64)
        la = np.zeros(n, np.int64)
        self._check(self.L.lemhip_profile_read(self.h, n, _ptr(ms), _ptr(la)), "profile_read")
        return {k: (float(ms[i]), int(la[i])) for i, k in enumerate(PROF_SLOTS)}

    def copy_to_last_climate(self) -> None:
        self._check(self.L.lemhip_copy_to_last_climate(self.h), "copy_to_last_climate")

    def tuning(self) -> dict:
        t = LemTuning()
        self._check(self.L.lemhip_get_tuning(self.h, C.byref(t)), "get_tuning")
        return t.as_dict()

    def set_tuning(self, dag_tile=None, dag_batch=None, plateau_rounds_per_launch=None, fp32_integer_fa=None,
                   use_graphs=None, fp32_tree_sediment=None, **_ignored) -> None:
        t = LemTuning()
        self._check(self.L.lemhip_get_tuning(self.h, C.byref(t)), "get_tuning")
        if fp32_tree_sediment is not None:
            t.fp32_tree_sediment = int(fp32_tree_sediment)
        if use_graphs is not None:
            t.use_graphs = int(use_graphs)
        if fp32_integer_fa is not None:
            t.fp32_integer_fa = int(fp32_integer_fa)
        if dag_tile is not None:
            t.dag_tile_w, t.dag_tile_h = dag_tile
        if dag_batch is not None:
            t.dag_batch = dag_batch
        if plateau_rounds_per_launch is not None:
            t.plateau_rounds_per_launch = plateau_rounds_per_launch
        self._check(self.L.lemhip_set_tuning(self.h, C.byref(t)), "set_tuning")

    def _snapshot(self, slot: int) -> bytes:
        n = C.c_size_t()
        self._check(self.L.lemhip_snapshot_size(self.h, slot, C.byref(n)), "snapshot_size")
        buf = np.empty(n.value, np.uint8)
        self._check(self.L.lemhip_snapshot(self.h, slot, _ptr(buf), n.value), "snapshot")
        return buf.tobytes()

    def _restore(self, slot: int, snap: bytes) -> None:
        buf = np.frombuffer(snap, np.uint8).copy()
        self._check(self.L.lemhip_restore(self.h, slot, _ptr(buf), buf.size), "restore")

    def autotune(self) -> None:
        """The next step times every launch candidate on this context's data
        and keeps the fastest (results do not depend on the choice)."""
        self._check(self.L.lemhip_autotune(self.h), "autotune")


class LemContext(_Handle):
    """One DEM with its state resident on the GPU."""

    def __init__(self, params: LemParams, precision: int = FP64, device: int = 0):
        self.L = lib()
        self.params = params
        self.precision = precision
        self.W, self.H = params.width, params.height
        self.N = self.W * self.H
        self.it = params.start_iter      # next iteration for step_group
        self._open(lambda ph: self.L.lemhip_create(C.byref(params), precision, device, ph), "lemhip_create")

    # --- state -----------------------------------------------------------
    def init_state(self, dem: np.ndarray, rain: np.ndarray | None = None, soilT_mode: str = "formula") -> None:
        z = np.ascontiguousarray(dem, np.float64).reshape(-1)
        assert z.size == self.N
        r = None if rain is None else np.ascontiguousarray(rain, np.float64).reshape(-1)
        mode = {"formula": 0, "nonneg": 1}[soilT_mode]
        self._check(self.L.lemhip_init_state(self.h, _ptr(z), None if r is None else _ptr(r), mode), "init_state")

    def download(self, name: str) -> np.ndarray:
        f = FIELDS[name]
        out = _buf(name, self.N)
        self._check(self.L.lemhip_download(self.h, f, _ptr(out)), f"download {name}")
        return out

    def upload(self, name: str, a: np.ndarray) -> None:
        f = FIELDS[name]
        dt = np.int32 if name in INT_FIELDS else np.float64
        a = np.ascontiguousarray(a, dt).reshape(-1)
        self._check(self.L.lemhip_upload(self.h, f, _ptr(a)), f"upload {name}")

    @property
    def outlet(self) -> int:
        o = C.c_int()
        self._check(self.L.lemhip_get_outlet(self.h, C.byref(o)), "get_outlet")
        return o.value

    @outlet.setter
    def outlet(self, v: int) -> None:
        self._check(self.L.lemhip_set_outlet(self.h, int(v)), "set_outlet")

    # --- phases (lem.cpp:83-97) -----------------------------------------
    def flow_direction(self, it: int) -> None:
        self._check(self.L.lemhip_flow_direction(self.h, it), "flow_direction")

    def runoff(self) -> None:
        self._check(self.L.lemhip_runoff(self.h), "runoff")

    def flow_accum(self, it: int) -> None:
        self._check(self.L.lemhip_flow_accum(self.h, it), "flow_accum")

    def erosion(self, it: int) -> None:
        self._check(self.L.lemhip_erosion(self.h, it), "erosion")

    def step(self, it: int) -> None:
        self._check(self.L.lemhip_step(self.h, it), "step")

    def run(self, it0: int, n: int) -> None:
        self._check(self.L.lemhip_run(self.h, it0, n), "run")

    def set_climate(self, rain: float, temp: float) -> None:
        self._check(self.L.lemhip_set_climate(self.h, rain, temp), "set_climate")

    def summary(self) -> dict:
        s = LemSummary()
        self._check(self.L.lemhip_summary(self.h, C.byref(s)), "summary")
        return s.as_dict()

    def snapshot(self) -> bytes:
        """Checkpoint of the resident state (lemhip_snapshot)."""
        return self._snapshot(0)

    def restore(self, snap: bytes) -> None:
        """Resume from a checkpoint of a DEM of this size and precision form."""
        self._restore(0, snap)


ATLAS_ALIGN = 32   # DEM rectangles start at columns and rows that are multiples of this


def atlas_span(sides) -> int:
    """Grid columns (or rows) that DEMs of the given widths (heights) take
    side by side in an atlas: each but the last is padded with its nodata
    column and rounded up to a multiple of ATLAS_ALIGN, the last needs its
    nodata column only."""
    sides = list(sides)
    if not sides:
        return 0
    a = ATLAS_ALIGN
    return sum(-(-(s + 1) // a) * a for s in sides[:-1]) + sides[-1] + 1


class LemAtlas(_Handle):
    """Many DEMs of different sizes, each in its own rectangle ("slot") of one
    device grid, stepped together.  Per-slot results equal those of a
    LemContext on the same inputs.  Rectangles are shelf-packed at columns
    and rows that are multiples of ATLAS_ALIGN (see atlas_span)."""

    MAX_SLOTS = 255

    def __init__(self, params: LemParams, width: int, height: int, precision: int = FP64, device: int = 0):
        self.L = lib()
        self.params = params
        self.precision = precision
        self.W, self.H = width, height
        self.iters = np.zeros(self.MAX_SLOTS, np.int32)
        self.shapes: dict[int, tuple[int, int]] = {}
        self._open(lambda ph: self.L.lemhip_create_atlas(C.byref(params), width, height, precision, device, ph),
                   "lemhip_create_atlas")

    def open(self, width: int, height: int) -> int | None:
        """Reserve a rectangle; None when the atlas has no room or no free slot."""
        s = C.c_int(-1)
        rc = self.L.lemhip_slot_open(self.h, width, height, C.byref(s))
        if rc == -3:
            return None
        self._check(rc, "slot_open")
        self.shapes[s.value] = (height, width)
        return s.value

    def init(self, slot: int, dem: np.ndarray, rain: np.ndarray | None = None, soilT_mode: str = "formula",
             iteration: int = -1000) -> None:
        h, w = self.shapes[slot]
        z = np.ascontiguousarray(dem, np.float64).reshape(-1)
        assert z.size == h * w
        r = None if rain is None else np.ascontiguousarray(rain, np.float64).reshape(-1)
        mode = {"formula": 0, "nonneg": 1}[soilT_mode]
        self._check(self.L.lemhip_slot_init(self.h, slot, _ptr(z), None if r is None else _ptr(r), mode),
                    "slot_init")
        self.iters[slot] = iteration

    def add(self, dem: np.ndarray, rain: np.ndarray | None = None, soilT_mode: str = "formula",
            iteration: int = -1000) -> int | None:
        dem = np.asarray(dem)
        s = self.open(dem.shape[1], dem.shape[0])
        if s is not None:
            self.init(s, dem, rain, soilT_mode, iteration)
        return s

    def remove(self, slot: int) -> None:
        self._check(self.L.lemhip_slot_close(self.h, slot), "slot_close")
        self.shapes.pop(slot, None)

    def rect(self, slot: int) -> tuple[int, int, int, int]:
        v = [C.c_int() for _ in range(4)]
        self._check(self.L.lemhip_slot_rect(self.h, slot, *[C.byref(x) for x in v]), "slot_rect")
        return tuple(x.value for x in v)

    def download(self, slot: int, name: str) -> np.ndarray:
        h, w = self.shapes[slot]
        out = _buf(name, h * w)
        self._check(self.L.lemhip_slot_download(self.h, slot, FIELDS[name], _ptr(out)), f"download {name}")
        return out

    def upload(self, slot: int, name: str, a: np.ndarray) -> None:
        dt = np.int32 if name in INT_FIELDS else np.float64
        a = np.ascontiguousarray(a, dt).reshape(-1)
        self._check(self.L.lemhip_slot_upload(self.h, slot, FIELDS[name], _ptr(a)), f"upload {name}")

    def outlet(self, slot: int) -> int:
        o = C.c_int()
        self._check(self.L.lemhip_slot_get_outlet(self.h, slot, C.byref(o)), "slot_get_outlet")
        return o.value

    def set_outlet(self, slot: int, v: int) -> None:
        self._check(self.L.lemhip_slot_set_outlet(self.h, slot, int(v)), "slot_set_outlet")

    def set_climate(self, slot: int, rain: float, temp: float) -> None:
        self._check(self.L.lemhip_slot_set_climate(self.h, slot, rain, temp), "slot_set_climate")

    def step(self) -> None:
        """One timestep of every active slot at its own iteration number."""
        self._check(self.L.lemhip_step_slots(self.h, _ptr(self.iters)), "step_slots")
        for s in self.shapes:
            self.iters[s] += 1

    def summary(self, slot: int) -> dict:
        s = LemSummary()
        self._check(self.L.lemhip_slot_summary(self.h, slot, C.byref(s)), "slot_summary")
        return s.as_dict()

    def snapshot(self, slot: int) -> bytes:
        return self._snapshot(slot)

    def restore(self, slot: int, snap: bytes, iteration: int) -> None:
        """Resume an open slot from a checkpoint; its next step is `iteration`."""
        self._restore(slot, snap)
        self.iters[slot] = iteration


def step_group(handles: list) -> None:
    """One timestep of several contexts at once (lemhip_step_group): each
    runs from its own host thread on its own stream, so their kernels overlap.
    Single-DEM contexts step at their `it` attribute (then incremented),
    atlases at their per-slot iterations."""
    L = lib()
    n = len(handles)
    hs = (C.c_void_p * n)(*[h.h.value for h in handles])
    its = (C.c_int * n)(*[getattr(h, "it", 0) for h in handles])
    IP = C.POINTER(C.c_int)
    slot_its = (IP * n)(*[h.iters.ctypes.data_as(IP) if isinstance(h, LemAtlas) else IP() for h in handles])
    rc = L.lemhip_step_group(hs, n, its, slot_its)
    if rc != 0:
        msgs = "; ".join(L.lemhip_last_error(h.h).decode() for h in handles)
        raise LemError(f"lemhip_step_group: {rc} {msgs}")
    for h in handles:
        if isinstance(h, LemAtlas):
            for s in h.shapes:
                h.iters[s] += 1
        else:
            h.it = getattr(h, "it", 0) + 1


# reference-named entry points (lem.cpp:84-96)
def cuFlowDirection(ctx: LemContext, iteration: int) -> None:
    ctx.flow_direction(iteration)


def computeRunOffWeights(ctx: LemContext) -> None:
    ctx.runoff()


def correctmfdflow(ctx: LemContext, iteration: int) -> None:
    ctx.flow_accum(iteration)


def erosionGPU(ctx: LemContext, iteration: int) -> None:
    ctx.erosion(iteration)


def math_eval(fn: int, x: np.ndarray, y: np.ndarray | None = None) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(x if y is None else y, np.float64)
    out = np.empty_like(x)
    rc = lib().lemhip_math_eval(fn, x.size, _ptr(x), _ptr(y), _ptr(out))
    if rc != 0:
        raise LemError(f"lemhip_math_eval: {rc}")
    return out


def mem_info(dev: int = 0) -> tuple[float, float]:
    """(free, total) device memory in bytes."""
    fr, to = C.c_double(), C.c_double()
    rc = lib().lemhip_mem_info(dev, C.byref(fr), C.byref(to))
    if rc != 0:
        raise LemError("lemhip_mem_info failed")
    return fr.value, to.value


def bytes_per_cell(precision: int, atlas: bool = True) -> tuple[float, float]:
    """(fixed, growth bound) device bytes per cell of a context
    (lemhip_bytes_per_cell)."""
    fx, gr = C.c_double(), C.c_double()
    rc = lib().lemhip_bytes_per_cell(precision, 1 if atlas else 0, C.byref(fx), C.byref(gr))
    if rc != 0:
        raise LemError("lemhip_bytes_per_cell failed")
    return fx.value, gr.value


def device_info(dev: int = 0) -> dict:
    name = C.create_string_buffer(64)
    cu = C.c_int()
    mem = C.c_double()
    rc = lib().lemhip_device_info(dev, name, 64, C.byref(cu), C.byref(mem))
    if rc != 0:
        raise LemError("lemhip_device_info failed")
    return {"arch": name.value.decode(), "cus": cu.value, "mem_gb": mem.value}

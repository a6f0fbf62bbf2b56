"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Imported by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
leg, and only as the checker.  Builds oracle/liblemoracle*.so on first use if
they are missing (gcc, no GPU).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class OrcParams(C.Structure):
    _fields_ = [
        ("W", C.c_int), ("H", C.c_int), ("cellsize", C.c_double), ("flowtype", C.c_int),
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


_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)


class OrcState(C.Structure):
    _fields_ = [
        ("dem", _DP), ("Slopes", _DP), ("SlopePtr", _DP),
        ("soilT", _DP), ("soilM", _DP), ("stone", _DP), ("fines", _DP), ("nut", _DP),
        ("soilB", _DP), ("TotB", _DP),
        ("rain", _DP), ("temp", _DP), ("last_rain", _DP), ("last_temp", _DP),
        ("mask", _IP), ("fd", _IP), ("SFD", _IP),
        ("w", _DP), ("fa", _DP), ("ero", _DP), ("inci", _DP), ("geli", _DP), ("depo", _DP),
        ("dz", _DP), ("wC", _DP), ("wP", _DP),
        ("outlet", C.c_int),
        ("FA_max", C.c_double), ("totE", C.c_double), ("totI", C.c_double), ("totG", C.c_double),
        ("totD", C.c_double), ("totwC", C.c_double), ("totwP", C.c_double), ("totbio", C.c_double),
        ("sinkcount", C.c_int), ("sinks_pass1", C.c_int),
        ("plateau_rounds1", C.c_int), ("plateau_rounds2", C.c_int),
        ("nlabels", C.c_int), ("nedges", C.c_int),
        ("fa_uncomputed", C.c_int), ("sed_uncomputed", C.c_int),
        ("totwater", C.c_double),
    ]


DOUBLE_FIELDS = ["dem", "SlopePtr", "soilT", "soilM", "stone", "fines", "nut", "soilB", "TotB",
                 "rain", "temp", "last_rain", "last_temp", "w", "fa", "ero", "inci", "geli",
                 "depo", "dz", "wC", "wP"]
INT_FIELDS = ["mask", "fd", "SFD"]

_libs: dict[bool, C.CDLL] = {}


def lib(openmp: bool = False) -> C.CDLL:
    if openmp in _libs:
        return _libs[openmp]
    name = "liblemoracle_omp.so" if openmp else "liblemoracle.so"
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        subprocess.check_call(["make", "-s", "-C", HERE])
    L = C.CDLL(path)
    L.orc_step.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcState), C.c_int]
    L.orc_flow_direction.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcState), C.c_int]
    L.orc_runoff.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcState)]
    L.orc_flow_accum.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcState), C.c_int]
    L.orc_erosion.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcState), C.c_int]
    L.orc_setProcessMatrices.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcState)]
    L.orc_SetClimateMatrices.argtypes = [C.POINTER(OrcParams), _IP, _DP, C.c_double, C.c_double, _DP, _DP]
    L.orc_init_setoutletcell.argtypes = [C.c_int, C.c_int, _IP, _DP]
    L.orc_createSoilTfromformula.argtypes = [C.c_int, _DP, _DP]
    L.orc_math_eval.argtypes = [C.c_int, C.c_int, _DP, _DP, _DP]
    L.orc_flow_accum_literal.argtypes = [C.c_int, C.c_int, _IP, _IP, _DP, _DP]
    L.orc_flow_accum_literal.restype = C.c_int
    L.orc_sediment_literal.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcState)]
    L.orc_sediment_literal.restype = C.c_int
    L.orc_flowdirs.argtypes = [C.c_int, C.c_int, C.c_int, _DP, _IP, _DP, _IP, _IP, _IP, _IP]
    L.orc_route_plateaus.argtypes = [C.c_int, C.c_int, _DP, _IP, _DP, _IP, _IP, C.POINTER(C.c_float)]
    L.orc_route_plateaus.restype = C.c_int
    L.orc_route_plateaus_literal.argtypes = L.orc_route_plateaus.argtypes
    L.orc_route_plateaus_literal.restype = C.c_int
    L.orc_watershed_labels.argtypes = [C.c_int, C.c_int, _DP, _IP, _IP, _IP]
    L.orc_watershed_labels.restype = C.c_int
    L.orc_spill_heights.argtypes = [C.c_int, C.c_int, _DP, _IP, _IP, C.c_int, _DP, _IP, _IP]
    L.orc_spill_heights.restype = C.c_int
    L.orc_set_threads.argtypes = [C.c_int]
    _libs[openmp] = L
    return L


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_IP)


def default_params(W: int, H: int, flowtype: int = 0, cellsize: float = 30.0,
                   runiniter: int = 10) -> OrcParams:
    """default.par:12-35 with the given grid."""
    return OrcParams(W=W, H=H, cellsize=cellsize, flowtype=flowtype,
                     stoneval=20.0, fineval=80.0, soilTval=1.0, soilMval=0.5, soilBval=0.0, nutval=10.0,
                     soil_struct=2, profile_permeability=3,
                     Qthreshold=6.0, Gslope=1.0, kappa=0.00001, Gflowlim=600.0,
                     Flow_A=0.00001, Flow_B=1.3, max_incision=0.001,
                     ddk=0.9, dck=0.4, dgk=0.9, temp_inflex=6.0, temp_sens=6.0,
                     rain=600.0, temp=10.0, runiniter=runiniter, start_iter=-1000)


class Oracle:
    """Host state of one DEM, stepped by the oracle."""

    def __init__(self, z: np.ndarray, rain: np.ndarray | None, params: OrcParams,
                 openmp: bool = False, soilT_mode: str = "formula"):
        self.L = lib(openmp)
        self.p = params
        H, W = z.shape
        assert (W, H) == (params.W, params.H)
        N = W * H
        self.N = N
        self.a = {f: np.zeros(N, np.float64) for f in DOUBLE_FIELDS}
        self.a.update({f: np.zeros(N, np.int32) for f in INT_FIELDS})
        self.a["Slopes"] = np.zeros(8 * N, np.float64)
        self.a["dem"][:] = z.reshape(-1)
        mask = self.a["mask"]
        mask[:] = (self.a["dem"] != -9999.0).astype(np.int32)
        s = OrcState()
        for f in DOUBLE_FIELDS + ["Slopes"]:
            setattr(s, f, dptr(self.a[f]))
        for f in INT_FIELDS:
            setattr(s, f, iptr(self.a[f]))
        self.s = s
        s.outlet = self.L.orc_init_setoutletcell(W, H, iptr(mask), dptr(self.a["dem"]))
        self.L.orc_createSoilTfromformula(N, dptr(self.a["dem"]), dptr(self.a["soilT"]))
        if soilT_mode == "nonneg":
            np.maximum(self.a["soilT"], 0.0, out=self.a["soilT"])
        elif soilT_mode != "formula":
            raise ValueError(soilT_mode)
        self.L.orc_SetClimateMatrices(C.byref(params), iptr(mask), dptr(self.a["dem"]),
                                      params.rain, params.temp, dptr(self.a["rain"]), dptr(self.a["temp"]))
        if rain is not None:
            r = rain.reshape(-1)
            self.a["rain"][:] = np.where(mask == 1, r, -9999.0)
        self.a["last_rain"][:] = self.a["rain"]
        self.a["last_temp"][:] = self.a["temp"]
        self.L.orc_setProcessMatrices(C.byref(params), C.byref(s))

    def step(self, it: int) -> None:
        self.L.orc_step(C.byref(self.p), C.byref(self.s), it)

    def __getitem__(self, k):
        return self.a[k]

    def scalars(self) -> dict:
        return {f: getattr(self.s, f) for f, _ in OrcState._fields_[26:]}


def math_eval(fn: int, x: np.ndarray, y: np.ndarray | None = None) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(x if y is None else y, np.float64)
    out = np.empty_like(x)
    lib().orc_math_eval(fn, x.size, dptr(x), dptr(y), dptr(out))
    return out

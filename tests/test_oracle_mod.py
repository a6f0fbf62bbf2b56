"""Pins the oracle's per-cell MOD formulas (diffuse and concentrated erosion,
gelifluction, weathering, surface change, nutrients, vegetation, the outlet
bookkeeping) with what the mathematics fixes rather than by retyping them:
scaling laws between cells that differ in one input, exact zeros and caps,
fixed points, and the published carbonate equilibrium constants.

Each test cell sits alone inside nodata, so sediment routing leaves it
untouched (every valid cell with a nodata neighbour to its W, E, SE, S or SW
is blocked, mfd_accum.cu:46,390-394) and its outputs depend on its own inputs
only.  Reference formulas: eroincidep.cu:39-198 (erosion, incision,
gelifluction), :200-327 (surface change, weathering), updates.cu:48-139
(nutrients, vegetation), updates.cu:142-181 (surface update, outlet).
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O
from parallem_amd import synth

NODATA = -9999.0
DEFAULT = dict(fa=0.05, s=0.2, stone=20.0, nut=10.0, TotB=100.0, fines=80.0, soilT=5.0, temp=10.0,
               rain=1000.0, soilB=5.0)


def run_cells(cells):
    """One isolated cell per entry; returns the oracle after one MOD pass."""
    K = len(cells)
    W, H = 2 * K + 1, 3
    z = np.full((H, W), NODATA)
    for k in range(K):
        z[1, 2 * k + 1] = 500.0
    o = O.Oracle(z, None, O.default_params(W, H), soilT_mode="nonneg")
    idx = []
    for k, c in enumerate(cells):
        v = dict(DEFAULT, **c)
        i = W + 2 * k + 1
        idx.append(i)
        o["fa"][i] = v["fa"]
        o["Slopes"][8 * i:8 * i + 8] = 0.0
        o["Slopes"][8 * i] = v["s"]
        for f in ("stone", "nut", "TotB", "fines", "soilT", "soilB", "rain"):
            o[f][i] = v[f]
        o["temp"][i] = v["temp"]
        o["last_temp"][i] = v["temp"]
        o["last_rain"][i] = v["rain"]
    before = {f: o[f].copy() for f in ("soilT", "soilB", "TotB", "dem")}
    o.L.orc_erosion(C.byref(o.p), C.byref(o.s), -1000)
    return o, idx, before


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_diffuse_erosion_scaling_laws():
    cells = [{}, {"fa": 0.10}, {"s": 0.4}, {"stone": 50.0}, {"TotB": 200.0}, {"nut": 12.0},
             {"fa": 1.0}, {"fa": 2.0}, {"fines": 40.0}, {"soilT": 1e-12}, {"s": 0.0}, {"s": 9.0}, {"s": 5.0}]
    o, i, _ = run_cells(cells)
    e = o["ero"][i]
    assert e[0] > 0
    assert rel(e[1] / e[0], 4.0) < 1e-12                       # OF^2, OF = fa*1000/25 below Qthreshold
    assert rel(e[2] / e[0], 2.0 ** 1.67) < 1e-12               # slope^1.67
    assert rel(e[3] / e[0], (60.0 / 90.0) ** 0.877) < 1e-12    # (110 - stone)^0.877
    assert rel(e[4] / e[0], math.exp(-0.07 * 100 / 76.5)) < 1e-12
    assert e[5] == 0.0                                         # (12 - nut) = 0, struct/permeability terms 0
    assert e[6] == e[7] and e[6] > e[1]                        # OF capped at Qthreshold = 6
    assert rel(e[8] / e[0], 0.5 ** 1.14) < 1e-12               # fines^1.14
    assert e[9] == 1e-12                                       # capped by soil thickness
    assert rel(e[10] / e[0], (1e-4 / 0.2) ** 1.67) < 1e-12     # zero slope -> 1e-4
    assert e[11] == e[12]                                      # slope capped at 5


def test_concentrated_erosion_threshold_and_exponent():
    # f = fa * cs^2/1e4 - Qthreshold = 0.09 fa - 6
    cells = [{"fa": 66.0}, {"fa": 100.0}, {"fa": 200.0}, {"fa": 100.0, "s": 0.9}, {"fa": 100.0, "s": 0.0},
             {"fa": 1e5}, {"fa": 100.0, "TotB": 200.0}]
    o, i, _ = run_cells(cells)
    c = o["inci"][i]
    assert c[0] == 0.0                                         # below the discharge threshold
    assert c[1] > 0
    assert rel(c[2] / c[1], (12.0 / 3.0) ** 1.3) < 1e-12       # f^Flow_B
    assert rel(c[3] / c[1], 0.5 / 0.2) < 1e-12                 # slope capped at 0.5, linear
    assert rel(c[4] / c[1], 0.01 / 0.2) < 1e-12                # zero slope -> 0.01
    assert c[5] == 0.001                                       # capped at max_incision
    assert rel(c[6] / c[1], math.exp(-0.07 * 100 / 76.5)) < 1e-12
    f = 0.09 * 100.0 - 6.0
    assert rel(c[1], 1e-5 * f ** 1.3 * 0.2 * math.exp(-0.07 * 100 / 76.5)) < 1e-12


def test_gelifluction_slope_flow_and_temperature():
    cells = [{"fa": 300.0}, {"fa": 300.0, "s": 0.4}, {"fa": 700.0}, {"fa": 300.0, "s": 2.0},
             {"fa": 0.05, "temp": 0.0}, {"fa": 150.0}]
    o, i, _ = run_cells(cells)
    g = o["geli"][i]
    assert rel(g[1] / g[0], 2.0) < 1e-12                       # linear in slope below Gslope
    assert g[2] == 0.0                                         # fa above Gflowlim: no gelifluction
    assert rel(g[3], 1e-5 * 1.0 * 0.5) < 1e-12                 # slope capped at Gslope = 1, par kappa above 8.87 C
    assert rel(g[4], 0.15711 * 0.2 * (1 - 0.05 / 600)) < 1e-12  # polynomial kappa at T = 0 is its constant term
    assert rel(g[5] / g[0], (1 - 150 / 600) / (1 - 300 / 600)) < 1e-12


def test_weathering_constants_match_plummer_busenberg():
    """log10(K1 KC KH / K2) at 25 C from the published constants
    (Plummer & Busenberg 1982: log K1 = -6.352, log K2 = -10.329,
    log KC = -8.480, log KH = -1.468) is -5.971; the reference's polynomial
    fits must reproduce it through the equilibrium Ca concentration."""
    T = 25.0
    fa = 10.0
    o, i, _ = run_cells([{"temp": T, "fa": fa, "soilT": 5.0}, {"temp": T, "fa": 2 * fa, "soilT": 5.0}])
    wC = o["wC"][i]
    assert rel(wC[1] / wC[0], 2.0) < 1e-12                     # proportional to surface flow
    rescale_w = 10000 * (10000 / 900.0)
    ca = wC[0] / (0.0148148148 * (fa * 1000) / rescale_w)
    pco2 = 10 ** (-2.5 + (T - 15) / 10)                        # soil thickness > 0.1
    dha = -0.4883 + 8.074e-4 * T
    dhb = -0.3241 + 1.600e-4 * T
    sq = math.sqrt(0.1)
    g_ca = 10 ** ((-4 * dha * sq) / (1 + 5e-8 * dhb * sq))
    g_hco3 = 10 ** ((-dha * sq) / (1 + 5.4e-8 * dhb * sq))
    logk = 3 * math.log10(ca) + math.log10(4 * g_ca * g_hco3 ** 2) - math.log10(pco2)
    assert abs(logk - (-5.971)) < 0.03, logk


def test_physical_weathering_and_surface_change():
    o, i, before = run_cells([{"soilT": 0.5}, {"soilT": 1.5}, {}])
    wP = o["wP"][i]
    assert rel(wP[1] / wP[0], math.exp(-2.3)) < 1e-12          # P0 exp(-P1 soilT)
    for k in i:
        assert o["dz"][k] == ((o["depo"][k] - o["ero"][k]) - o["inci"][k]) - o["geli"][k]
        assert o["depo"][k] == 1e-5                             # blocked cell keeps the initial deposition
    # soil thickness gains the weathering products (no sediment update on blocked cells)
    k = i[2]
    assert o["soilT"][k] == before["soilT"][k] + o["wC"][k] * 0.02 + o["wP"][k]


def test_vegetation_fixed_point_and_nutrients_on_bare_rock():
    o, i, before = run_cells([{}, {"rain": 400.0, "temp": -5.0}, {"soilT": -1.0, "soilB": 5.0},
                              {"rain": 0.0, "temp": 30.0}])
    for k, (rain, T) in zip(i[:2], ((1000.0, 10.0), (400.0, -5.0))):
        cc = (3000.0 + 5.0 * (rain - (100 + 100 / (T + 15)))) / (10000 / 900.0)
        assert rel(o["TotB"][k], cc) < 1e-12                    # temperature unchanged: biomass at capacity
    assert o["TotB"][i[3]] >= 1.0                               # floor
    k = i[2]
    assert o["soilT"][k] == 0.0 and o["stone"][k] == 0.0 and o["soilM"][k] == 0.02
    assert o["soilB"][k] > 0 and o["nut"][k] == 100.0           # bare rock with biomass: nutrient-saturated


def test_outlet_drops_by_0_00024_per_step():
    dem = synth.fbm_512(side=64, seed_offset=21)
    o = O.Oracle(dem.z, dem.rain, O.default_params(64, 64), soilT_mode="nonneg")
    out = o.s.outlet
    z0 = o["dem"][out]
    m = o["mask"] == 1
    for k in range(3):
        o.step(-1000 + k)
        assert o.s.outlet == out
        assert abs(o["dem"][out] - (z0 - 0.00024 * (k + 1))) < 1e-9
        base = o["dem"][out] + 0.00012
        others = m.copy()
        others[out] = False
        assert np.all(o["dem"][others] > base)                  # floored at base + 0.01

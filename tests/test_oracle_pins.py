"""Pins of the oracle's remaining update rules by what the mathematics fixes:
conservation laws, exact zeros, limits, fixed points, ratios between cells
that differ in one input, and hand grids whose answer is known.

* runoff: water balance w + d(soilM) = rain/1000 below the 0.5 cap, the cap,
  and the runoff coefficient's stone factor (runoffweight.cu:21-58);
* weathering: stone does not move without erosion, moves down when it does,
  and its relative change follows tanh(ero/soilT) for small ratios; the soil
  moisture drawdown is 1e-3 * ET / soilT with ET from the PE comparison
  (eroincidep.cu:290-324);
* nutrients: leaf fall TotB / (72.8 ln(TotB + 1.4)) at TotB = e - 1.4,
  decomposition linear in soilB with rate exp(0.12 T - 5.75) = 1 at
  T = 5.75/0.12, nutrient % proportional to soilB / soilT (updates.cu:48-76);
* vegetation with a temperature change: the carrying capacity is a fixed
  point of logistic growth, and at T = temp_inflex the rates are exactly dT
  (updates.cu:78-139);
* the MFD fd/SFD merge on a pit filled by flooding, fd[outlet] = 16
  (routeflow.cu:938-953), and the outlet check (routeflow.cu:171-184,
  739-741).

Each isolated test cell sits inside nodata, so sediment routing leaves it
untouched (mfd_accum.cu:46,390-394) and its outputs depend on its own inputs.
"""
import ctypes as C
import math

import numpy as np

from oracle import oracle as O
from parallem_amd import synth

NODATA = -9999.0
DEFAULT = dict(fa=0.05, s=0.2, stone=20.0, nut=10.0, TotB=100.0, fines=80.0, soilT=5.0, temp=10.0,
               rain=1000.0, soilB=5.0, soilM=0.5)


def isolated(cells):
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
        for f in ("stone", "nut", "TotB", "fines", "soilT", "soilB", "rain", "soilM"):
            o[f][i] = v[f]
        o["temp"][i] = v["temp"]
        o["last_temp"][i] = v.get("last_temp", v["temp"])
        o["last_rain"][i] = v["rain"]
    return o, idx


def erosion(cells):
    o, idx = isolated(cells)
    before = {f: o[f][idx].copy() for f in ("stone", "soilM", "soilB", "TotB", "soilT")}
    o.L.orc_erosion(C.byref(o.p), C.byref(o.s), -1000)
    return o, idx, before


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


# ----------------------------------------------------------------- runoff

def test_runoff_water_balance_and_cap():
    # light rain keeps soil moisture below its 0.5 cap in cells 0-2, 4, 5
    cells = [{"soilM": 0.1, "rain": 200.0}, {"soilM": 0.1, "rain": 300.0},
             {"soilM": 0.02, "stone": 60.0, "rain": 200.0}, {"soilM": 0.45, "rain": 1000.0},
             {"soilM": 0.1, "stone": 60.0, "rain": 200.0}, {"soilM": 0.3, "rain": 200.0}]
    o, i = isolated(cells)
    m0 = o["soilM"][i].copy()
    o.L.orc_runoff(C.byref(o.p), C.byref(o.s))
    w, m1 = o["w"][i], o["soilM"][i]
    for k in (0, 1, 2, 4, 5):
        # rain (mm/yr) / 1000 splits exactly into runoff and soil moisture
        rain = cells[k]["rain"] / 1000
        assert m1[k] < 0.5
        assert abs(w[k] + (m1[k] - m0[k]) - rain) < 1e-15
        assert 0 < w[k] < rain
    assert rel(w[1] / w[0], 1.5) < 1e-14                       # linear in rain
    assert m1[3] == 0.5                                        # capped: the excess is lost, not runoff
    assert w[3] + (0.5 - m0[3]) < 1.0
    # the stone factor (1 - 0.0085 stone) at equal moisture and biomass
    assert rel(w[4] / w[0], (1 - 0.0085 * 60) / (1 - 0.0085 * 20)) < 1e-13
    # the moisture factor tanh(0.6 + soilM/0.2) at equal stone and biomass
    assert rel(w[5] / w[0], math.tanh(0.6 + 1.5) / math.tanh(0.6 + 0.5)) < 1e-13
    # nodata cells carry -9999 and are outside the water total
    assert o["w"][0] == -9999
    assert rel(o.s.totwater, float(np.sum(w))) < 1e-15


# ------------------------------------------------------------- weathering

def test_weathering_stone_needs_erosion_and_moves_down():
    # nut = 12 zeroes diffuse erodibility (12 - nut = 0 and the structure and
    # permeability terms are 0 at the default par values); fa = 0.05 stays
    # below the incision threshold, so ero = inci = 0
    cells = [{"nut": 12.0}, {}, {"s": 0.4}, {"s": 0.1}]
    o, i, b = erosion(cells)
    assert o["ero"][i[0]] == 0.0 and o["inci"][i[0]] == 0.0
    assert o["stone"][i[0]] == b["stone"][0]                   # no erosion: stone unchanged
    assert o["fines"][i[0]] == 100.0 - b["stone"][0]
    for k in (1, 2, 3):
        assert o["ero"][i[k]] > 0
        assert o["stone"][i[k]] < b["stone"][k]                 # chemical weathering removes stone
        assert o["fines"][i[k]] == 100.0 - o["stone"][i[k]]
    # small ero/soilT: the relative stone change is linear in the erosion
    # (tanh x = x + O(x^3)); cells 1-3 differ only in slope, so their
    # erosion ratios follow slope^1.67 and so must their stone changes
    d = [(b["stone"][k] - o["stone"][i[k]]) / b["stone"][k] for k in (1, 2, 3)]
    e = [o["ero"][i[k]] for k in (1, 2, 3)]
    assert rel(e[1] / e[0], 2.0 ** 1.67) < 1e-12
    assert rel(d[1] / d[0], e[1] / e[0]) < 1e-6
    assert rel(d[2] / d[0], e[2] / e[0]) < 1e-6


def test_weathering_soil_moisture_drawdown():
    # rain 1000 + surface flow 1000*fa = 1050 mm against PE = 1.44 (T + 6)^2:
    # T = 10: PE = 368.64 < 1050, ET = 1050 - 368.64; T = 25: PE = 1383.84,
    # ET = 1050
    cells = [{"soilT": 5.0}, {"soilT": 10.0}, {"temp": 25.0, "soilT": 5.0}]
    o, i, b = erosion(cells)
    sT = o["soilT"][i]
    draw = b["soilM"] - o["soilM"][i]
    assert rel(draw[0] * sT[0], 1e-3 * (1050.0 - 1.44 * 16 ** 2)) < 1e-12
    assert rel(draw[1] * sT[1], 1e-3 * (1050.0 - 1.44 * 16 ** 2)) < 1e-12
    assert rel(draw[2] * sT[2], 1e-3 * 1050.0) < 1e-12
    # inversely proportional to the (post-weathering) soil thickness
    assert rel(draw[0] / draw[1], sT[1] / sT[0]) < 1e-12


# -------------------------------------------------------------- nutrients

def test_leaf_fall_decomposition_and_nutrient_percent():
    e = math.e
    T0 = 5.75 / 0.12                          # exp(0.12 T - 5.75) = 1
    cells = [{"TotB": e - 1.4, "soilB": 0.0},
             {"TotB": e - 1.4, "soilB": 4.0, "temp": T0},
             {"TotB": e - 1.4, "soilB": 8.0, "temp": T0},
             {"TotB": e - 1.4, "soilB": 4.0},
             {"TotB": e - 1.4, "soilB": 8.0}]
    o, i, b = erosion(cells)
    leaf = (e - 1.4) / 72.8                   # natural log: ln(e) = 1
    sb = o["soilB"][i]
    assert rel(sb[0], leaf) < 1e-14
    assert abs(sb[1] - leaf) < 1e-14          # decomposition = soilB at rate 1
    assert abs(sb[2] - leaf) < 1e-14
    # linear in soilB at any temperature
    assert rel((b["soilB"][4] + leaf - sb[4]) / (b["soilB"][3] + leaf - sb[3]), 2.0) < 1e-12
    assert rel((b["soilB"][3] + leaf - sb[3]) / 4.0, math.exp(0.12 * 10.0 - 5.75)) < 1e-12
    # nutrient percent = soilB / (1300 soilT) * 100
    sT = o["soilT"][i]
    for k in range(5):
        assert rel(o["nut"][i[k]], sb[k] / (1300.0 * sT[k]) * 100.0) < 1e-14


# ------------------------------------------------------------- vegetation

def test_vegetation_logistic_with_temperature_change():
    T = 6.0                                   # temp_inflex: tanh term 0, rates equal dT
    # capacity at (T, rain) from the no-change branch (TotB = cc, pinned by
    # test_vegetation_fixed_point_and_nutrients_on_bare_rock)
    o, i, _ = erosion([{"temp": T}])
    cc = float(o["TotB"][i[0]])
    cells = [{"temp": T, "last_temp": T - 0.5, "TotB": cc},        # fixed point of growth
             {"temp": T, "last_temp": T - 0.5, "TotB": cc / 2},    # +dT * tb * (1 - 1/2)
             {"temp": T, "last_temp": T + 0.5, "TotB": cc},        # -dT * cc * (cc/cc)
             {"temp": T, "last_temp": T + 0.5, "TotB": cc / 4},
             {"temp": T, "last_temp": T + 2.0, "TotB": cc}]        # floor at 1
    o, i, _ = erosion(cells)
    tb = o["TotB"][i]
    assert tb[0] == cc
    assert rel(tb[1], cc / 2 + 0.5 * (cc / 2) * 0.5) < 1e-14
    assert rel(tb[2], cc - 0.5 * cc) < 1e-14
    assert rel(tb[3], cc / 4 - 0.5 * (cc / 4) * 0.25) < 1e-14
    assert tb[4] == 1.0


# ----------------------------------------------------- merge and outlet

def test_mfd_merge_gives_a_filled_pit_its_raised_surface_direction():
    # a plane falling towards row 0 (nodata frame around it) with a pit at
    # (3, 4): in MFD mode every other cell keeps its multi-bit FlowDirs code;
    # the pit has no downhill neighbour, so the merge gives it the SFD of the
    # raised surface, where it is filled to the lowest of its neighbours and
    # routed onto that neighbour
    H, W = 8, 9
    z = np.full((H, W), NODATA)
    for r in range(1, H - 1):
        for c in range(1, W - 1):
            z[r, c] = 10.0 * r + 0.5 * abs(c - 4)
    z[3, 4] = 15.0                              # pit (the outlet stays at (1, 4), 10 m)
    p = O.default_params(W, H, flowtype=1)
    o = O.Oracle(z, None, p, soilT_mode="nonneg")
    o.L.orc_flow_direction(C.byref(p), C.byref(o.s), -1000)
    fd = o["fd"].reshape(H, W)
    # neighbours of the pit: the lowest is (2, 4) at 20 m, due north (code 64)
    assert fd[3, 4] == 64
    # a cell with several downhill neighbours keeps all of them: (5, 2) at
    # 51 m has N (41), NE (40.5), E (50.5) and NW (41.5) below it
    assert fd[5, 2] == 64 | 128 | 32 | 1
    # fd[outlet] = 16 (routeflow.cu:953)
    assert o["fd"][o.s.outlet] == 16


def test_outlet_check_lowers_a_raised_outlet():
    dem = synth.fbm_512(side=64, seed_offset=9)
    o = O.Oracle(dem.z, dem.rain, O.default_params(64, 64), soilT_mode="nonneg")
    o.step(-1000)
    valid = np.flatnonzero(o["mask"] == 1)
    pick = int(valid[len(valid) // 2])
    lowest = float(o["dem"][valid].min())
    assert lowest < o["dem"][pick]
    o.s.outlet = pick
    p = o.p
    o.L.orc_flow_direction(C.byref(p), C.byref(o.s), -999)
    # lowered after FlowDirs: the directions saw the old height, the
    # surface now holds lowest - 1e-4 there
    assert o["dem"][pick] == lowest - 1e-4
    assert o["fd"][pick] == 16

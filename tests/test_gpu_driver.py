"""The C++ driver program (parallem_amd/driver, lem.cpp:14-128 over
include/lemhip_compat.h and the C-ABI) on a DEM written as an ESRI ASCII
grid with a 35-line par file: after 3 steps its fields are bit-identical
with the CPU oracle, its budgets CSV carries the oracle's totals, the
periodic grids are written, and a run resumed from its own checkpoint ends
where the unbroken run ends."""
import os
import subprocess

import numpy as np
import pytest

from parallem_amd import synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "parallem_amd", "driver", "lem_driver")

PAR = [  # default.par's values, in its line order
    "0", "0", "1", "10", "{flow}", "600.0", "10.0", "20.0", "80.0", "2.0", "3.0", "1.0", "0.5", "0.0", "10.0",
    "6.0", "1.0", "0.00001", "600.0", "0.00001", "1.3", "0.001", "0.0", "0.0", "0.9", "0.4", "0.9", "6.0", "6.0"]


def setup_run(tmp_path, dem, flow):
    H, W = dem.z.shape
    with open(tmp_path / "dem.asc", "w") as f:
        f.write(f"ncols {W}\nnrows {H}\nxllcorner 0\nyllcorner 0\ncellsize 30\nNODATA_value -9999\n")
        for r in range(H):
            f.write(" ".join(repr(float(v)) for v in dem.z[r]) + "\n")
    (tmp_path / "clim.txt").write_text("year ppt temp\n0 600 10\n")
    lines = ["driver test", "03/12/2026", "run", "dem.asc", "clim.txt", "none.tif"]
    lines += [v.format(flow=flow) for v in PAR]
    (tmp_path / "run.par").write_text("\n".join(f"{x:<40}#" for x in lines) + "\n")


def run(tmp_path, *args):
    print(f"[driver] {' '.join(args)}", flush=True)
    r = subprocess.run([DRIVER, "run.par", *args], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r


def load(d, name, dt):
    return np.fromfile(os.path.join(d, name), dtype=dt)


@pytest.mark.parametrize("flow", [0, 1])
def test_driver_matches_the_oracle(tmp_path, flow):
    from oracle import oracle as O
    dem = synth.fbm_512(side=96, seed_offset=51)
    H, W = dem.z.shape
    setup_run(tmp_path, dem, flow)
    run(tmp_path, "--steps", "3", "--soilT-nonneg", "--dump", "out")
    # the driver's rain is the par file's gradient (SetClimateMatrices), so
    # the oracle starts from the same rain field
    orc = O.Oracle(dem.z, None, O.default_params(W, H, flow), soilT_mode="nonneg")
    rows = []
    for k in range(3):
        orc.step(-1000 + k)
        rows.append(orc.scalars())
    out = tmp_path / "out"
    for name, dt, ref in (("dem.f64", np.float64, orc["dem"]), ("fa.f64", np.float64, orc["fa"]),
                          ("soilT.f64", np.float64, orc["soilT"]), ("stone.f64", np.float64, orc["stone"])):
        assert np.array_equal(load(out, name, dt), ref), name
    fd = load(out, "fd.i32", np.int32)
    assert np.array_equal(fd, orc["fd"])
    csv = (tmp_path / "run" / "run_budgets.csv").read_text().splitlines()
    assert len(csv) == 4
    for k, line in enumerate(csv[1:]):
        v = [x.strip() for x in line.split(",")]
        assert int(v[0]) == -1000 + k
        assert float(v[4]) == pytest.approx(rows[k]["totE"], rel=1e-6, abs=1e-8)
        assert float(v[9]) == pytest.approx(rows[k]["totD"], rel=1e-6, abs=1e-8)
        assert float(v[11]) == pytest.approx(rows[k]["FA_max"], abs=1e-3)
        assert int(v[14]) == rows[k]["sinkcount"]
    # iteration -1000 is a multiple of 50 and 100: all seven grids written
    for g in ("FD", "FA", "height", "ero", "inci", "geli", "depo"):
        assert (tmp_path / "run" / f"-1000_{g}.asc").exists(), g
    h = np.loadtxt(tmp_path / "run" / "-1000_height.asc", skiprows=6)
    assert h.shape == (H, W)
    log = (tmp_path / "run" / "run_logfile.txt").read_text()
    assert "Total amount of water in grid for runoff" in log


def test_driver_restart_resumes_bit_for_bit(tmp_path):
    dem = synth.fbm_512(side=80, seed_offset=52)
    setup_run(tmp_path, dem, 0)
    run(tmp_path, "--steps", "4", "--fp32", "--soilT-nonneg", "--checkpoint-every", "2", "--dump", "full")
    assert (tmp_path / "run" / "restart" / "-999_state.lemsnap").exists()
    run(tmp_path, "--steps", "4", "--fp32", "--soilT-nonneg", "--restart", "-998", "--dump", "resumed")
    for name, dt in (("dem.f64", np.float64), ("fa.f64", np.float64), ("soilT.f64", np.float64),
                     ("stone.f64", np.float64), ("fd.i32", np.int32)):
        a, b = load(tmp_path / "full", name, dt), load(tmp_path / "resumed", name, dt)
        assert np.array_equal(a, b, equal_nan=dt == np.float64), name

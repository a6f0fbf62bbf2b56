"""The reference driver's file formats (io.cpp), as built into the C++ driver
(parallem_amd/driver/lem_io.cpp), through liblemio.so: the 35-line
positional parameter file (io.cpp:376-580), ESRI ASCII grids (io.cpp:28-199),
the climate series (io.cpp:265-319) and the budgets CSV (io.cpp:233-263).
No GPU needed."""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "parallem_amd", "driver", "liblemio.so")


@pytest.fixture(scope="module")
def io():
    if not os.path.exists(LIB):
        import subprocess
        subprocess.check_call(["make", "-s", "-C", os.path.dirname(LIB), "liblemio.so"])
    L = C.CDLL(LIB)
    dp, ip, cp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_char_p
    L.lemio_read_par.argtypes = [cp, dp, cp, cp, cp, cp]
    L.lemio_read_grid.argtypes = [cp, ip, dp, dp, C.c_long]
    L.lemio_write_grid.argtypes = [cp, C.c_int, C.c_int, C.c_int, dp, ip, C.c_void_p]
    L.lemio_write_summary.argtypes = [cp, C.c_int, C.c_int, C.c_double, C.c_double, dp, C.c_int]
    L.lemio_read_climate.argtypes = [cp, C.c_int, ip, dp, dp]
    return L


# the order of the positional lines (after title and date): names, then
# numbers in the order readinpar assigns them
PAR_VALUES = [0, -5, 40, 7, 1, 750.0, 8.5, 25.0, 75.0, 2, 3, 1.5, 0.45, 0.25, 11.0, 5.5, 0.9, 2e-5,
              550.0, 2e-5, 1.25, 0.002, 0.0, 0.0, 0.85, 0.35, 0.8, 7.0, 5.0]


def write_par(path, names=("run42", "dem.asc", "clim.txt", "rock.tif"), values=PAR_VALUES):
    lines = ["a parameter file written by the tests", "03/12/2026"]
    for n in names:
        lines.append(f"{n:<40}# a file or run name")
    for v in values:
        lines.append(f"{v!s:<40}# value")
    path.write_text("\n".join(lines) + "\n")


def test_par_file_positions_and_names(io, tmp_path):
    p = tmp_path / "run.par"
    write_par(p)
    out = (C.c_double * 29)()
    names = [C.create_string_buffer(256) for _ in range(4)]
    assert io.lemio_read_par(str(p).encode(), out, *names) == 35
    assert [n.value.decode() for n in names] == ["run42", "dem.asc", "clim.txt", "rock.tif"]
    assert list(out) == [float(v) for v in PAR_VALUES]


def test_par_file_blanks_and_truncation(io, tmp_path):
    # blanks inside a name are dropped (stripblank); only the first 38
    # characters of a line count, so a comment starting after column 38 is
    # never part of the name; numbers read from the first 10 (integers) or
    # 20 (reals) characters
    p = tmp_path / "run.par"
    write_par(p, names=("my run", "dem file.asc", "c.txt", "b.tif"))
    out = (C.c_double * 29)()
    names = [C.create_string_buffer(256) for _ in range(4)]
    assert io.lemio_read_par(str(p).encode(), out, *names) == 35
    assert names[0].value == b"myrun" and names[1].value == b"demfile.asc"
    short = tmp_path / "short.par"
    short.write_text("title\n" * 20)
    assert io.lemio_read_par(str(short).encode(), out, *names) < 0
    assert io.lemio_read_par(str(tmp_path / "missing.par").encode(), out, *names) < 0


def test_esri_grid_round_trip(io, tmp_path):
    rng = np.random.default_rng(4)
    h, w = 7, 9
    z = rng.uniform(0, 800, (h, w)).round(6)
    mask = np.ones((h, w), np.int32)
    mask[0, :] = 0
    mask[:, -1] = 0
    hdr = np.array([707862.5, 4467835.25, 30.0, -9999.0])
    f = tmp_path / "g.asc"
    assert io.lemio_write_grid(str(f).encode(), 0, w, h, hdr.ctypes.data_as(C.POINTER(C.c_double)),
                               mask.ctypes.data_as(C.POINTER(C.c_int)), z.ctypes.data) == 0
    text = f.read_text().splitlines()
    assert text[0] == f"ncols {w}" and text[1] == f"nrows {h}"
    assert text[5] == "NODATA_value -9999.000000"
    assert text[6].split()[0] == "-9999"                        # masked cells as the integer -9999
    wh = (C.c_int * 2)()
    hdr4 = (C.c_double * 4)()
    vals = np.zeros(h * w)
    assert io.lemio_read_grid(str(f).encode(), wh, hdr4, vals.ctypes.data_as(C.POINTER(C.c_double)), h * w) == 0
    assert (wh[0], wh[1]) == (w, h) and list(hdr4) == list(hdr)
    got = vals.reshape(h, w)
    assert np.array_equal(got[mask == 1], z[mask == 1])          # %lf keeps 6 decimals
    assert np.all(got[mask == 0] == -9999)
    # integer grids
    fd = rng.integers(0, 255, (h, w)).astype(np.int32)
    g2 = tmp_path / "fd.asc"
    assert io.lemio_write_grid(str(g2).encode(), 1, w, h, hdr.ctypes.data_as(C.POINTER(C.c_double)),
                               mask.ctypes.data_as(C.POINTER(C.c_int)), fd.ctypes.data) == 0
    assert io.lemio_read_grid(str(g2).encode(), wh, hdr4, vals.ctypes.data_as(C.POINTER(C.c_double)), h * w) == 0
    assert np.array_equal(vals.reshape(h, w)[mask == 1], fd[mask == 1])


def test_esri_grid_header_spellings_and_short_file(io, tmp_path):
    f = tmp_path / "t.asc"
    f.write_text("NCOLS 3\nNROWS 2\nXLLCENTER 1.5\nYLLCENTER 2.5\nCELLSIZE 10\nnodata_value -1\n"
                 "1 2 3\n4 5 6\n")
    wh = (C.c_int * 2)()
    hdr4 = (C.c_double * 4)()
    vals = np.zeros(6)
    assert io.lemio_read_grid(str(f).encode(), wh, hdr4, vals.ctypes.data_as(C.POINTER(C.c_double)), 6) == 0
    assert list(vals) == [1, 2, 3, 4, 5, 6] and list(hdr4) == [1.5, 2.5, 10.0, -1.0]
    f.write_text("ncols 3\nnrows 2\nxllcorner 0\nyllcorner 0\ncellsize 30\nNODATA_value -9999\n1 2 3\n")
    assert io.lemio_read_grid(str(f).encode(), wh, hdr4, vals.ctypes.data_as(C.POINTER(C.c_double)), 6) != 0
    assert io.lemio_read_grid(str(tmp_path / "none.asc").encode(), wh, hdr4, None, 0) != 0


def test_budgets_csv_lines(io, tmp_path):
    f = tmp_path / "b.csv"
    tot = np.array([1.5, 2.25, 0.125, 0.001, 0.002, 0.5, 1234.5, 987.654321])
    tp = tot.ctypes.data_as(C.POINTER(C.c_double))
    assert io.lemio_write_summary(str(f).encode(), -1000, -1000, 10.0, 600.0, tp, 7) == 0
    assert io.lemio_write_summary(str(f).encode(), -999, -1000, 10.0, 600.0, tp, 8) == 0
    lines = f.read_text().splitlines()
    assert len(lines) == 3
    head = [x.strip() for x in lines[0].split(",")]
    assert head == ["iteration", "year", "temp", "rain", "diffuse", "concentrated", "gelifluction", "chemicalW",
                    "physicalW", "deposition", "totbio", "FA_max", "sedoutflow", "waterout", "sinkcount"]
    row = [x.strip() for x in lines[2].split(",")]
    assert row[0] == "-999" and row[1] == "-999" and row[14] == "8"
    assert row[2] == "10.000000" and row[3] == "600.000000"
    assert row[4] == "1.50000000"                                  # %10.8lf
    assert float(row[5]) == 2.25 and len(row[5].split(".")[1]) == 81   # the reference's %10.81lf
    assert row[11] == "987.654"                                     # %10.3lf
    assert float(row[12]) == pytest.approx((1.5 + 2.25 + 0.125 - 0.5) / 1300000, abs=1e-8)
    assert row[13] == "987654.32100"                                # FA_max * 1000, %10.5lf


def test_climate_series(io, tmp_path):
    f = tmp_path / "clim.txt"
    f.write_text("year ppt temp\n0 650.5 9.5\n1 700.25 10.25\n2 710 11\n")
    y = (C.c_int * 3)()
    r = (C.c_double * 3)()
    t = (C.c_double * 3)()
    assert io.lemio_read_climate(str(f).encode(), 3, y, r, t) == 0
    assert list(y) == [0, 1, 2] and list(r) == [650.5, 700.25, 710.0] and list(t) == [9.5, 10.25, 11.0]
    assert io.lemio_read_climate(str(f).encode(), 5, y, r, t) != 0     # fewer rows than iterations

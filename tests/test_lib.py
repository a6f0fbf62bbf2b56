"""The C-ABI library builds, loads without a GPU, and exports every entry
point that include/lemhip.h declares (no compute calls here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h") or fn == "lem_math.h":
            continue
        src = open(os.path.join(ROOT, "include", fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\*?\s*(lemhip_\w+)\s*\(", src, flags=re.M)
    return names


def test_header_declares_the_entry_points():
    names = declared()
    for n in ("lemhip_create", "lemhip_destroy", "lemhip_flow_direction", "lemhip_runoff",
              "lemhip_flow_accum", "lemhip_erosion", "lemhip_step", "lemhip_summary"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from parallem_amd import lemhip
    L = lemhip.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    assert set(lemhip.EXPORTS) <= set(declared())


def test_no_cpu_fallback_without_library(tmp_path, monkeypatch):
    from parallem_amd import lemhip
    monkeypatch.setattr(lemhip, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(lemhip, "_lib", None)
    try:
        lemhip.lib()
    except lemhip.LemError:
        return
    raise AssertionError("loading a missing library must raise")


def test_bad_arguments_are_status_codes_not_crashes():
    from parallem_amd import lemhip
    L = lemhip.lib()
    p = lemhip.default_params(0, 10)
    h = ctypes.c_void_p()
    assert L.lemhip_create(ctypes.byref(p), 0, 0, ctypes.byref(h)) == -1
    p = lemhip.default_params(10, 10)
    assert L.lemhip_create(ctypes.byref(p), 7, 0, ctypes.byref(h)) == -1
    assert L.lemhip_step(None, 0) == -1


def test_atlas_span_rounds_rectangles_to_the_alignment():
    """Rectangles start at multiples of ATLAS_ALIGN: each DEM but the last
    takes its width plus the nodata column rounded up, the last only its
    width plus the nodata column."""
    from parallem_amd import lemhip
    assert lemhip.ATLAS_ALIGN == 32
    assert lemhip.atlas_span([]) == 0
    assert lemhip.atlas_span([100]) == 101
    assert lemhip.atlas_span([4096, 4096]) == 4128 + 4097
    assert lemhip.atlas_span([31, 31, 31]) == 32 + 32 + 32
    assert lemhip.atlas_span([32, 5]) == 64 + 6


def test_summary_and_profile_layouts_match_the_header():
    """The Python mirrors of lem_summary and the profile slots end where the
    header's do (a field added on one side only would shift every later one)."""
    from parallem_amd import lemhip
    hdr = open(os.path.join(ROOT, "include", "lemhip.h")).read()
    body = hdr[hdr.index("typedef struct {\n    double FA_max"):hdr.index("} lem_summary;")]
    last = [l.strip() for l in body.splitlines() if l.strip().startswith("int ") or l.strip().startswith("double ")][-1]
    assert last == "int sed_open;"
    assert lemhip.LemSummary._fields_[-1][0] == "sed_open"
    assert len(lemhip.PROF_SLOTS) == 16 and lemhip.PROF_SLOTS[-1] == "sed_contract"

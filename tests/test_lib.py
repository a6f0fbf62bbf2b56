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

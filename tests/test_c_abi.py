"""The C ABI library loads and exports every function include/cpab_hip.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = []
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(cpab_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_api():
    names = declared()
    for must in ("cpab_create", "cpab_calc_T", "cpab_prep_trels", "cpab_calc_v",
                 "cpab_calc_cell_idx", "cpab_calc_trajectory", "cpab_remap2d", "cpab_ssd_ll"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from cpabdiffeo_amd.build import SO
    assert os.path.exists(SO), "run python -m cpabdiffeo_amd.build first"
    lib = ctypes.CDLL(SO)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    from cpabdiffeo_amd import _lib
    assert sorted(_lib.SIGNATURES) == declared()


def test_version_string():
    from cpabdiffeo_amd import _lib
    assert b"gfx950" in _lib.lib().cpab_version()

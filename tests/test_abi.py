"""The C-ABI library loads and exports every function include/elegant_hip.h
declares (no GPU work: nothing is called)."""
import ctypes as C
import os
import re

from elegant_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "elegant_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return sorted(set(re.findall(r"\b(eh_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared()
    for n in ("eh_multipole_tracking2", "eh_track_through_csbend", "eh_simple_rf_cavity", "eh_exactDrift",
              "eh_drift_beam", "eh_limit_amplitudes", "eh_removeInvalidParticles", "eh_do_tracking"):
        assert n in names


def test_library_exports_every_declared_function():
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_loader_list_matches_header():
    assert sorted(_lib.EXPORTS) == declared()

"""CPU checks of the C ABI: the library builds/loads and exports every
entry point include/fcg.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

from fermicg_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "fcg.h")).read()
    return sorted(set(re.findall(r"\b(fcg_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "run __graft_entry__.build()"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared()
    assert len(names) >= 19
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTS)


def test_binding_signatures_load():
    L = _lib.load()
    assert L.fcg_last_error.restype is ctypes.c_char_p

"""The C-ABI boundary: the shared library loads without a GPU and exports
every function include/egsgpu.h declares (no compute calls here)."""
import ctypes
import os
import re

from egsnrc_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "egsgpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(egsgpu_\w+)\s*\(", txt)))


def test_library_exports_every_declared_function():
    names = declared()
    assert len(names) > 20 and "egsgpu_create" in names and "egsgpu_set_tallies" in names
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing

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


def _err(L):
    buf = ctypes.create_string_buffer(1024)
    L.egsgpu_last_error(buf, 1024)
    return buf.value.decode()


def test_argument_errors_are_returned_not_fatal():
    """The boundary returns a status and keeps a message (no exit()): null
    handles and arguments, an unreadable table file, and the source and
    geometry checks egsgpu_create makes before it touches a device."""
    import numpy as np
    from egsnrc_amd.engine import Tables, make_fano_source, make_source
    L = _lib.lib()
    E_ARG = -1
    assert L.egsgpu_run(None, ctypes.c_int64(10)) == E_ARG
    assert "invalid argument" in _err(L)
    assert L.egsgpu_set_streams(None, 0, 0) == E_ARG
    n = ctypes.c_int32(0)
    assert L.egsgpu_graph_stats(None, ctypes.byref(n), ctypes.byref(n), ctypes.byref(n)) == E_ARG
    t = ctypes.c_void_p()
    rc = L.egsgpu_tables_load(b"/nonexistent/tables.egstab", ctypes.byref(t))
    assert rc < 0 and rc != E_ARG and "nonexistent" in _err(L)
    tab = Tables(os.path.join(ROOT, "egsnrc_amd", "data", "521icru.egstab"))
    xb = np.linspace(-1.0, 1.0, 5)
    g, keep = _lib.make_geometry(xb, xb.copy(), xb.copy(), np.zeros(64, np.int32), None)
    tp = _lib.default_transport()
    h = ctypes.c_void_p()

    def create(src):
        return L.egsgpu_create(tab.h, ctypes.byref(g), ctypes.byref(src[0]),
                               ctypes.byref(tp), ctypes.byref(h))
    assert L.egsgpu_create(tab.h, None, None, None, ctypes.byref(h)) == E_ARG
    bad = make_source(2, 1.0)                      # no such charge
    assert create(bad) == E_ARG and "charge" in _err(L)
    bad = make_source(0, 1.0)
    bad[0].kind = 99
    assert create(bad) == E_ARG and "source kind" in _err(L)
    fano = make_fano_source(-1, 1.0, (-1, 1, -1, 1, 1, -1), rho_max=1.0)   # empty z range
    assert create(fano) == E_ARG and "Fano" in _err(L)
    g.nx = 0
    assert create(make_source(0, 1.0)) == E_ARG and "geometry" in _err(L)

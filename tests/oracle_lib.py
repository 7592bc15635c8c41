"""Loader for the CPU oracle (oracle/liboracle.so) used as the checker.

The oracle is test infrastructure: only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg load it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_LIB):
        build()
    L = C.CDLL(ORACLE_LIB)
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    for name, args in {
        "oracle_ranmar_fill": [i32, i32, i32, i64, vp],
        "oracle_ranlux_fill": [i32, i32, i32, i64, vp],
        "oracle_trace_rays": [vp, i32, vp, vp, i32, i32, vp, vp, vp],
        "oracle_hownear": [vp, i32, vp, vp, i32, vp],
        "oracle_is_where": [vp, i32, vp, i32, vp],
    }.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def ranmar_fill(ixx, jxx, nstreams, n):
    out = np.zeros(nstreams * n)
    assert load().oracle_ranmar_fill(ixx, jxx, nstreams, n, _p(out)) == 0
    return out.reshape(nstreams, n)


def ranlux_fill(lux, seed, nstreams, n):
    out = np.zeros(nstreams * n)
    assert load().oracle_ranlux_fill(lux, seed, nstreams, n, _p(out)) == 0
    return out.reshape(nstreams, n)


def trace_rays(geom, x, u, max_steps, precision):
    x = np.ascontiguousarray(x, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    n = x.shape[0]
    reg = np.zeros((n, max_steps), dtype=np.int32)
    st = np.zeros((n, max_steps))
    ns = np.zeros(n, dtype=np.int32)
    assert load().oracle_trace_rays(C.byref(geom), n, _p(x), _p(u), max_steps,
                                    precision, _p(reg), _p(st), _p(ns)) == 0
    return reg, st, ns


def hownear(geom, x, ireg, precision):
    x = np.ascontiguousarray(x, dtype=np.float64)
    ireg = np.ascontiguousarray(ireg, dtype=np.int32)
    out = np.zeros(x.shape[0])
    assert load().oracle_hownear(C.byref(geom), x.shape[0], _p(x), _p(ireg),
                                 precision, _p(out)) == 0
    return out


def is_where(geom, x, precision):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(x.shape[0], dtype=np.int32)
    assert load().oracle_is_where(C.byref(geom), x.shape[0], _p(x), precision,
                                  _p(out)) == 0
    return out

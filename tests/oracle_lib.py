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


class OracleSimulation:
    """The oracle's transport run with the same inputs as
    egsnrc_amd.engine.Simulation (checker only)."""

    def __init__(self, blob_path, geom, source, transport, nstreams=1024,
                 nthreads=8):
        L = load()
        for name, args in {
            "oracle_create": [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
            "oracle_set_streams": [C.c_void_p, C.c_int32, C.c_int32],
            "oracle_seed": [C.c_void_p, C.c_int32, C.c_int32],
            "oracle_run": [C.c_void_p, C.c_int64],
            "oracle_get_tallies": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
            "oracle_destroy": [C.c_void_p],
            "oracle_last_error": [C.c_void_p, C.c_size_t],
        }.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        self.L = L
        self.geom = geom
        self.nvox = geom.nx * geom.ny * geom.nz
        self.h = C.c_void_p()
        self._check(L.oracle_create(blob_path.encode(), C.byref(geom),
                                    C.byref(source), C.byref(transport),
                                    C.byref(self.h)))
        self._check(L.oracle_set_streams(self.h, nstreams, nthreads))

    def _check(self, rc):
        if rc != 0:
            buf = C.create_string_buffer(1024)
            self.L.oracle_last_error(buf, 1024)
            raise RuntimeError("oracle: " + buf.value.decode())

    def seed(self, ixx, jxx):
        self._check(self.L.oracle_seed(self.h, ixx, jxx))

    def run(self, n):
        self._check(self.L.oracle_run(self.h, int(n)))

    def tallies(self):
        s = np.zeros(self.nvox)
        s2 = np.zeros(self.nvox)
        n = C.c_int64(0)
        self._check(self.L.oracle_get_tallies(self.h, _p(s), _p(s2), C.byref(n)))
        return s, s2, n.value

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.oracle_destroy(self.h)
            self.h = C.c_void_p()


def sample(blob, kind, medium, energy, n, param=0.0, seed=9373):
    """Run one interaction sampler of the oracle n times (oracle_sample)."""
    L = load()
    f = L.oracle_sample
    f.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_double, C.c_double,
                  C.c_int64, C.c_int32, C.c_void_p]
    f.restype = C.c_int
    out = np.zeros((n, 4))
    assert f(blob.encode(), kind, medium, energy, param, n, seed, _p(out)) == 0
    return out


def grad(sim, n, target, mode=0):
    """Score-function density gradient of the oracle (oracle_grad).  mode 0:
    L = sum (D - D*)^2; mode 1: total energy deposited per history."""
    L = sim.L
    f = L.oracle_grad
    f.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_int32]
    f.restype = C.c_int
    target = np.ascontiguousarray(target, dtype=np.float64)
    g = np.zeros(sim.nvox)
    d = np.zeros(sim.nvox)
    loss = C.c_double(0)
    sim._check(f(sim.h, int(n), _p(target), _p(g), _p(d), C.byref(loss), mode))
    return g, d, loss.value

"""ctypes bindings of the C ABI in include/egsgpu.h.

The engine library ``libegsgpu.so`` is built in-tree by
``egsnrc_amd/build.py``.  Loading fails loudly when it is missing: there is
no CPU fallback for any entry point.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libegsgpu.so")

_lib = None


class EgsXYZGeometry(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("xb", C.c_void_p), ("yb", C.c_void_p), ("zb", C.c_void_p),
                ("medium", C.c_void_p), ("rhor", C.c_void_p),
                ("boundary_tolerance", C.c_double)]


class EgsSource(C.Structure):
    _fields_ = [("charge", C.c_int32), ("spectrum", C.c_int32),
                ("energy", C.c_double), ("nbins", C.c_int32),
                ("spec_e", C.c_void_p), ("spec_w", C.c_void_p),
                ("shape", C.c_int32),
                ("x0", C.c_double), ("x1", C.c_double), ("y0", C.c_double),
                ("y1", C.c_double), ("z", C.c_double),
                ("ux", C.c_double), ("uy", C.c_double), ("uz", C.c_double),
                ("kind", C.c_int32), ("rot", C.c_double * 9), ("iso", C.c_double * 3),
                ("xcol", C.c_double), ("ycol", C.c_double), ("dsource", C.c_double),
                ("z1", C.c_double), ("rho_max", C.c_double)]


class EgsTransport(C.Structure):
    _fields_ = [("ecut", C.c_double), ("pcut", C.c_double),
                ("rayleigh", C.c_int32), ("bound_compton", C.c_int32),
                ("photoelectron_angle", C.c_int32),
                ("atomic_relaxation", C.c_int32), ("pair_angle", C.c_int32),
                ("brems_angle", C.c_int32), ("spin", C.c_int32),
                ("exact_bca", C.c_int32), ("estepe", C.c_double),
                ("ximax", C.c_double), ("skin_depth", C.c_double),
                ("kerma", C.c_int32), ("fp64", C.c_int32), ("howfarless", C.c_int32),
                ("n_split", C.c_int32), ("nbr_split", C.c_int32), ("play_rr", C.c_int32),
                ("ireject", C.c_int32), ("esave", C.c_double),
                ("step_algorithm", C.c_int32), ("brems_xsec", C.c_int32),
                ("pair_xsec", C.c_int32), ("photon_xsec", C.c_char * 16),
                ("eii_xsec", C.c_char * 16)]


def default_transport(**kw):
    """EGSnrc defaults (egsnrc.macros:2253-2330) with spin effects off."""
    t = EgsTransport(ecut=0.0, pcut=0.0, rayleigh=1, bound_compton=3,
                     photoelectron_angle=1, atomic_relaxation=1, pair_angle=1,
                     brems_angle=1, spin=0, exact_bca=1, estepe=0.25,
                     ximax=0.5, skin_depth=3.0, kerma=0, fp64=0, howfarless=0,
                     n_split=0, nbr_split=0, play_rr=0, ireject=0, esave=0.0,
                     step_algorithm=0, brems_xsec=0, pair_xsec=0)
    for k, v in kw.items():
        setattr(t, k, v)
    return t


SCHED_HISTORY = 0
SCHED_EVENT = 1
SCHED_HISTORY_PLAIN = 2
RNG_RANMAR = 0
RNG_RANLUX = 1
SOURCE_PARALLEL_BEAM = 0
SOURCE_DOSXYZ_FRONT = 1
SOURCE_DOSXYZ_BEAM = 2
SOURCE_PHSP = 3
SOURCE_DOSXYZ_POINT = 4
SOURCE_FANO = 5


class EngineError(RuntimeError):
    pass


def lib():
    """Load libegsgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise EngineError("libegsgpu.so is not built; run python -m egsnrc_amd.build")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    sig = {
        "egsgpu_last_error": [vp, C.c_size_t],
        "egsgpu_device_count": [vp],
        "egsgpu_ranmar_fill": [i32, i32, i32, i64, vp],
        "egsgpu_ranlux_fill": [i32, i32, i32, i64, vp],
        "egsgpu_trace_rays": [vp, i32, vp, vp, i32, i32, vp, vp, vp],
        "egsgpu_hownear": [vp, i32, vp, vp, i32, vp],
        "egsgpu_atomic_rate": [i32, i32, i64, vp, vp],
        "egsgpu_tables_load": [C.c_char_p, vp],
        "egsgpu_tables_free": [vp],
        "egsgpu_tables_nmedia": [vp, vp],
        "egsgpu_create": [vp, vp, vp, vp, vp],
        "egsgpu_destroy": [vp],
        "egsgpu_set_streams": [vp, i32, i32],
        "egsgpu_seed": [vp, i32, i32],
        "egsgpu_set_rng": [vp, i32, i32, i32],
        "egsgpu_set_phsp": [vp, i64, vp, vp, vp, vp, vp, vp, vp],
        "egsgpu_run": [vp, i64],
        "egsgpu_get_tallies": [vp, vp, vp, vp],
        "egsgpu_reset_tallies": [vp],
        "egsgpu_set_tallies": [vp, vp, vp, C.c_int64],
        "egsgpu_get_counters": [vp, vp],
        "egsgpu_get_energy_balance": [vp, vp],
        "egsgpu_run_static": [vp, i64],
        "egsgpu_set_photon_tracking": [vp, i32, C.c_double],
        "egsgpu_last_kernel_ms": [vp, vp, vp],
        "egsgpu_density_grad": [vp, i64, vp, vp, vp, vp],
        "egsgpu_tables_medium_info": [vp, i32, vp, C.c_size_t, vp],
        "egsgpu_set_schedule": [vp, i32],
        "egsgpu_set_event_shape": [vp, i32, i32, i32],
        "egsgpu_set_event_batching": [vp, i32, i32, i32, i32],
        "egsgpu_kernel_time": [vp, i32, vp, vp],
        "egsgpu_kernel_counters": [vp, i32, vp],
        "egsgpu_checkpoint": [vp, vp, vp],
        "egsgpu_run_batched": [vp, i32, vp],
        "egsgpu_autotune": [vp, i64, vp],
        "egsgpu_set_event_graph": [vp, i32],
        "egsgpu_graph_stats": [vp, vp, vp, vp],
        "egsgpu_restore": [vp, vp, C.c_size_t],
    }
    for name, args in sig.items():
        f = getattr(L, name, None)
        if f is None:
            continue
        f.argtypes = args
        f.restype = C.c_int
    _lib = L
    return L


def exported_symbols():
    """Names declared in include/egsgpu.h that the library exports."""
    L = lib()
    import re
    hdr = open(os.path.join(HERE, "..", "include", "egsgpu.h")).read()
    names = sorted(set(re.findall(r"\bint\s+(egsgpu_\w+)\s*\(", hdr)))
    return {n: hasattr(L, n) for n in names}


def check(rc):
    if rc != 0:
        buf = C.create_string_buffer(1024)
        lib().egsgpu_last_error(buf, 1024)
        raise EngineError("egsgpu error %d: %s" % (rc, buf.value.decode()))


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def make_geometry(xb, yb, zb, medium=None, rhor=None, tol=0.0):
    """Build an EgsXYZGeometry; returns (struct, keepalive list)."""
    xb = np.ascontiguousarray(xb, dtype=np.float64)
    yb = np.ascontiguousarray(yb, dtype=np.float64)
    zb = np.ascontiguousarray(zb, dtype=np.float64)
    nx, ny, nz = len(xb) - 1, len(yb) - 1, len(zb) - 1
    keep = [xb, yb, zb]
    med = None
    if medium is not None:
        med = np.ascontiguousarray(medium, dtype=np.int32).reshape(-1)
        keep.append(med)
    rh = None
    if rhor is not None:
        rh = np.ascontiguousarray(rhor, dtype=np.float64).reshape(-1)
        keep.append(rh)
    g = EgsXYZGeometry(nx, ny, nz, ptr(xb).value, ptr(yb).value, ptr(zb).value,
                       None if med is None else ptr(med).value,
                       None if rh is None else ptr(rh).value, tol)
    return g, keep


def ranmar_fill(ixx, jxx, nstreams, n):
    out = np.zeros(nstreams * n)
    check(lib().egsgpu_ranmar_fill(ixx, jxx, nstreams, n, ptr(out)))
    return out.reshape(nstreams, n)


def ranlux_fill(lux, seed, nstreams, n):
    out = np.zeros(nstreams * n)
    check(lib().egsgpu_ranlux_fill(lux, seed, nstreams, n, ptr(out)))
    return out.reshape(nstreams, n)


def trace_rays(geom, x, u, max_steps, precision):
    x = np.ascontiguousarray(x, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    n = x.shape[0]
    reg = np.zeros((n, max_steps), dtype=np.int32)
    st = np.zeros((n, max_steps))
    ns = np.zeros(n, dtype=np.int32)
    check(lib().egsgpu_trace_rays(C.byref(geom), n, ptr(x), ptr(u), max_steps,
                                  precision, ptr(reg), ptr(st), ptr(ns)))
    return reg, st, ns


def hownear(geom, x, ireg, precision):
    x = np.ascontiguousarray(x, dtype=np.float64)
    ireg = np.ascontiguousarray(ireg, dtype=np.int32)
    out = np.zeros(x.shape[0])
    check(lib().egsgpu_hownear(C.byref(geom), x.shape[0], ptr(x), ptr(ireg),
                               precision, ptr(out)))
    return out


def atomic_rate(pattern, naddr, nops):
    """FP64 atomic adds per second of egsgpu_atomic_rate, and the number of
    adds the table holds afterwards (both launches)."""
    ms = C.c_double(0)
    total = C.c_double(0)
    check(lib().egsgpu_atomic_rate(pattern, naddr, int(nops), C.byref(ms), C.byref(total)))
    per = 64
    issued = ((int(nops) + per - 1) // per + 255) // 256 * 256 * per
    return issued / (ms.value * 1e-3), total.value, issued

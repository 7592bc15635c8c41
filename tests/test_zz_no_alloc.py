"""egsgpu_run_static (the run that allocates nothing, its history tables
and stacks in the library's static device array) against egsgpu_run on
the same simulation: the same histories with the same random numbers give
the same event counters and the same dose sums up to the order of the FP64
additions, on two calls in a row, the second smaller than the first."""
import ctypes as C
import os

import numpy as np
import pytest

from egsnrc_amd import _lib, workloads
from egsnrc_amd.engine import Simulation, Tables, make_source

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BLOB521 = os.path.join(ROOT, "egsnrc_amd", "data", "521icru.egstab")


def kerma():
    xb, yb, zb, med, rhor = workloads.kerma_phantom(32, 0.8)
    src = make_source(0, 2.0, x0=-4, x1=4, y0=-4, y1=4, z=-1.0)
    return xb, yb, zb, med, rhor, src, _lib.default_transport(ecut=20.0, kerma=1)


def coupled():
    n, dx = 24, 0.5
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    zc = 0.5 * (zb[1:] + zb[:-1])
    mz = np.select([zc < 3, zc < 6, zc < 8], [0, 1, 2], 0)
    med = np.broadcast_to(mz[:, None, None], (n, n, n)).reshape(-1).astype(np.int32)
    e, w = workloads.falling_spectrum()
    src = make_source(0, 0.0, spectrum=(e, w), x0=-4, x1=4, y0=-4, y1=4, z=-1.0)
    return xb, xb.copy(), zb, med, None, src, _lib.default_transport(ecut=0.521, pcut=0.01)


def run_static(sim, n):
    _lib.check(_lib.lib().egsgpu_run_static(sim.h, C.c_int64(n)))


@pytest.mark.gpu
@pytest.mark.parametrize("case,schedule,streams,fp64,sizes", [
    ("kerma", "history", 8192, 0, (300000, 70001)),
    ("kerma", "history-plain", 4096, 1, (200000, 50000)),
    ("coupled", "history-plain", 1024, 0, (20000, 5003)),
])
def test_run_static_matches_run(case, schedule, streams, fp64, sizes):
    xb, yb, zb, med, rhor, src, tp = {"kerma": kerma, "coupled": coupled}[case]()
    tp.fp64 = fp64
    tab = Tables(BLOB521)
    a = Simulation(tab, xb, yb, zb, med, rhor, src, tp, nstreams=streams, schedule=schedule)
    b = Simulation(tab, xb, yb, zb, med, rhor, src, tp, nstreams=streams, schedule=schedule)
    for s in (a, b):
        s.run(0)   # sets up the streams (and is the same for both)
    total = 0
    for n in sizes:
        a.run(n)
        run_static(b, n)
        total += n
        ta, tb = a.tallies(), b.tallies()
        ca, cb = a.counters(), b.counters()
        print("%s %s fp64=%d after %d: counters %s" % (case, schedule, fp64, total, list(cb)))
        assert ta[2] == tb[2] == total
        assert list(ca) == list(cb)
        assert cb[2] > 0 and cb[5] == 0 and cb[6] == 0
        np.testing.assert_allclose(tb[0], ta[0], rtol=1e-12, atol=0)
        np.testing.assert_allclose(tb[1], ta[1], rtol=1e-12, atol=0)
        ea, eb = a.energy_balance(), b.energy_balance()
        for k in ("incident", "escaped"):
            assert eb[k] == pytest.approx(ea[k], rel=1e-12)


@pytest.mark.gpu
def test_run_static_limits():
    xb, yb, zb, med, rhor, src, tp = coupled()
    tab = Tables(BLOB521)
    big = Simulation(tab, xb, yb, zb, med, rhor, src, tp, nstreams=4096,
                     schedule="history-plain")
    big.run(0)
    with pytest.raises(_lib.EngineError, match="scratch"):
        run_static(big, 1000)
    fresh = Simulation(tab, xb, yb, zb, med, rhor, src, tp, nstreams=1024,
                       schedule="history-plain")
    with pytest.raises(_lib.EngineError, match="streams"):
        run_static(fresh, 1000)
    ev = Simulation(tab, xb, yb, zb, med, rhor, src, tp, nstreams=1024, schedule="event")
    ev.run(0)
    with pytest.raises(_lib.EngineError, match="history schedules"):
        run_static(ev, 1000)

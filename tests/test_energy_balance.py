"""Energy conservation of the shower, end to end on the device: the energy
brought into the grid by the primaries equals the energy deposited in the
voxels plus the energy carried out of the grid (EGSnrc hands every
particle's energy to ausgab when it is deposited or discarded,
egsnrc.mortran:1773-1816, 6600-6643).  This pins the engine without the
oracle: a dropped particle, a sampler that loses or creates energy, a
positron whose annihilation is forgotten or a double-counted deposit all
break it.  FP64 transport balances to rounding (1e-9); FP32 transport to the
rounding of its energy arithmetic (1e-5)."""
import os

import numpy as np
import pytest

from egsnrc_amd import _lib, workloads
from egsnrc_amd.engine import Simulation, Tables, make_source

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BLOB521 = os.path.join(ROOT, "egsnrc_amd", "data", "521icru.egstab")


def slabs(n=24, dx=0.5):
    """water / lung / bone / water slabs along z, n^3 voxels of dx cm"""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    zc = 0.5 * (zb[1:] + zb[:-1])
    mz = np.select([zc < 3, zc < 6, zc < 8], [0, 1, 2], 0)
    med = np.broadcast_to(mz[:, None, None], (n, n, n)).reshape(-1).astype(np.int32)
    return xb, xb.copy(), zb, med, None


CASES = {
    # name: (charge, energy or spectrum, transport keywords)
    "photon spectrum": (0, None, dict(ecut=0.521, pcut=0.01)),
    "6 MeV photons, pair": (0, 6.0, dict(ecut=0.521, pcut=0.01)),
    "10 MeV electrons": (-1, 10.0, dict(ecut=0.521, pcut=0.01)),
    "8 MeV positrons": (1, 8.0, dict(ecut=0.521, pcut=0.01)),
    "2 MeV photons, kerma": (0, 2.0, dict(ecut=20.0, pcut=0.01, kerma=1)),
}


def run(name, schedule, fp64, nhist):
    charge, energy, kw = CASES[name]
    xb, yb, zb, med, rhor = slabs()
    # a beam wider than the phantom: particles enter the sides too
    if energy is None:
        e, w = workloads.falling_spectrum()
        src = make_source(charge, 0.0, spectrum=(e, w), x0=-4, x1=4, y0=-4, y1=4, z=-1.0)
    else:
        src = make_source(charge, energy, x0=-4, x1=4, y0=-4, y1=4, z=-1.0)
    tp = _lib.default_transport(fp64=fp64, **kw)
    sim = Simulation(Tables(BLOB521), xb, yb, zb, med, rhor, src, tp, nstreams=8192,
                     schedule=schedule)
    sim.run(nhist // 2)
    sim.run(nhist - nhist // 2)
    b = sim.energy_balance()
    c = sim.counters()
    sim.close()
    return b, c


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("schedule,fp64", [("history", 1), ("history", 0), ("event", 0)])
def test_incident_equals_deposited_plus_escaped(name, schedule, fp64):
    if schedule == "event" and CASES[name][2].get("kerma"):
        schedule = "history-plain"
    b, c = run(name, schedule, fp64, 100000)
    inc, dep, esc = b["incident"], b["deposited"], b["escaped"]
    print("%s %s fp64=%d: incident %.6f deposited %.6f escaped %.6f (rel. residual %.2e)"
          % (name, schedule, fp64, inc, dep, esc, (inc - dep - esc) / inc))
    assert c[5] == 0 and c[6] == 0
    charge, energy, _ = CASES[name]
    if energy is not None:
        # every primary enters the grid with its kinetic energy (+2 m c^2
        # for a positron)
        expect = energy + (2 * 0.5109989461 if charge > 0 else 0.0)
        assert abs(inc - expect) < 1e-6 * expect
    assert dep > 0.05 * inc and esc > 0.01 * inc
    tol = 1e-9 if fp64 else 1e-5
    assert abs(inc - dep - esc) < tol * inc


@pytest.mark.gpu
def test_balance_survives_checkpoint_and_reset():
    """The bookkeeping regions are kept by a checkpoint and cleared by a
    reset, like the dose."""
    b1, _ = run("10 MeV electrons", "history", 1, 20000)
    charge, energy, kw = CASES["10 MeV electrons"]
    xb, yb, zb, med, rhor = slabs()
    src = make_source(charge, energy, x0=-4, x1=4, y0=-4, y1=4, z=-1.0)
    tab = Tables(BLOB521)
    a = Simulation(tab, xb, yb, zb, med, rhor, src, _lib.default_transport(fp64=1, **kw),
                   nstreams=4096, schedule="history")
    a.run(20000)
    ck = a.checkpoint()
    ba = a.energy_balance()
    b = Simulation(tab, xb, yb, zb, med, rhor, src, _lib.default_transport(fp64=1, **kw),
                   nstreams=4096, schedule="history")
    b.restore(ck)
    bb = b.energy_balance()
    for k in ("incident", "escaped", "deposited"):
        assert bb[k] == pytest.approx(ba[k], rel=1e-15)
    _lib.check(_lib.lib().egsgpu_reset_tallies(b.h))
    b.run(1000)
    bc = b.energy_balance()
    assert bc["incident"] == pytest.approx(energy, rel=1e-12)

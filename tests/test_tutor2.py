"""tutor2 (user_codes/tutor2/tutor2.mortran) as a reference-held end-to-end
check: 20 MeV electrons at normal incidence on a 1 mm tantalum plate
(tutor_data.pegs4dat, ECUT 1.5 MeV, PCUT 0.1 MeV), as a one-voxel phantom
2000 cm wide.  The fraction of the energy deposited in the plate is held to
the two runs the reference's documentation prints (tests/golden/
tutor2_fractions.json) within 3 of their standard errors, which follow
from the per-history spread of the deposit and their history counts
(10000 and 1000), plus 3 % of the value: both reference runs had spin
effects on (the engine runs without: spinms.data is not in the reference
tree), which add large-angle scattering in tantalum and so path length in
the plate, and the 2003 run predates the XCOM photon data, EADL relaxation
and bound Compton defaults.  The check is loose by design: it catches a
transport that is wrong by more than that, not the spin effect itself."""
import json
import os

import numpy as np
import pytest

from egsnrc_amd import _lib
from egsnrc_amd.engine import Simulation, Tables, make_source
from tests import oracle_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TUTOR = os.path.join(ROOT, "egsnrc_amd", "data", "tutor.egstab")
REF = json.load(open(os.path.join(ROOT, "tests", "golden", "tutor2_fractions.json")))
E0 = 20.0


def plate():
    xb = np.array([-1000.0, 1000.0])
    zb = np.array([0.0, 0.1])
    med = np.zeros(1, dtype=np.int32)          # TA
    src = make_source(-1, E0, shape="point", x0=0.0, y0=0.0, z=-1.0)
    tp = _lib.default_transport(ecut=1.5, pcut=0.1, kerma=0)
    return xb, xb.copy(), zb, med, None, src, tp


def judge(t, label):
    s, s2, n = t
    r = s[0] / n
    sd = np.sqrt(max(s2[0] / n - r * r, 0.0))          # per history
    dep = 100 * r / E0
    for run in REF["runs"]:
        se = 100 * sd / E0 / np.sqrt(run["histories"])
        print("%s: deposited %.3f %% vs %s %.3f %% (+- %.3f)"
              % (label, dep, run["label"], run["deposited"], se))
        assert abs(dep - run["deposited"]) < 3 * se + 0.03 * run["deposited"]
    return dep, 100 * sd / E0 / np.sqrt(n)


def oracle_run(n=1_000_000):
    xb, yb, zb, med, rhor, src, tp = plate()
    g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
    o = oracle_lib.OracleSimulation(TUTOR, g, src[0], tp, nstreams=256, nthreads=8)
    o.run(n)
    return o.tallies()


def test_oracle_tutor2_deposit():
    judge(oracle_run(), "oracle")


@pytest.mark.gpu
def test_engine_tutor2_deposit(engine):
    """The engine against the reference's runs, and against the oracle
    within 4 combined standard errors."""
    xb, yb, zb, med, rhor, src, tp = plate()
    sim = Simulation(Tables(TUTOR), xb, yb, zb, med, rhor, src, tp)
    sim.run(10_000_000)
    d_e, se_e = judge(sim.tallies(), "engine")
    d_o, se_o = judge(oracle_run(), "oracle")
    assert abs(d_e - d_o) < 4 * np.hypot(se_e, se_o), (d_e, d_o)
    eb = sim.energy_balance()
    assert abs(eb["incident"] - E0) < 1e-9 * E0

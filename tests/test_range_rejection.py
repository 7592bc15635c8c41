"""DOSXYZnrc charged-particle range rejection (IREJECT, ESAVE_GLOBAL;
$USER-RANGE-DISCARD at dosxyznrc.mortran:1050-1064 with RANGE_ECUTRR from
:2514-2545): below ESAVE_GLOBAL an electron or positron whose range down to
ECUT does not reach the nearest voxel boundary deposits its kinetic energy
in its voxel at once.

The rejection drops only the bremsstrahlung such an electron would have
made below ESAVE_GLOBAL, a small fraction of a per cent of its energy in
water at the energies used here, so the dose must agree with the oracle's
unrejected transport within statistics while the number of
condensed-history steps drops."""
import os

import numpy as np
import pytest

from egsnrc_amd import _lib
from egsnrc_amd.engine import Simulation, Tables, make_source
from egsnrc_amd import workloads
from tests import oracle_lib
from tests.parity import check_parity
from tests.test_dosxyz_frontend import parse, write_isource1_input

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BLOB521 = os.path.join(ROOT, "egsnrc_amd", "data", "521icru.egstab")

REC13 = "100000, 0, 500.0, 33, 97, 100.0, 0, 0, 0, 0, , 0, 0, 0, 0, 0, 0"


def transport_options(p):
    t = p["transport"]
    return dict(zip(t[::2], t[1::2]))


@pytest.mark.parametrize("rec,ireject,esave,note", [
    ("100000, 0, 500.0, 33, 97, 100.0, 0, 0, 0, 1, 2.0, 0, 0, 0, 0, 0, 0", "1", "2", None),
    ("100000, 0, 500.0, 33, 97, 100.0, 0, 0, 0, 1, 0.6, 0, 0, 0, 0, 0, 0", "0", "0",
     "ESAVE_GLOBAL <= ECUT"),
    (REC13, "0", "0", None)])
def test_parse_range_rejection(tmp_path, rec, ireject, esave, note):
    """IREJECT and ESAVE_GLOBAL (record 13, fields 10-11); no rejection when
    ESAVE_GLOBAL <= ECUT (ECUT 0.7 here), as DOSXYZnrc decides."""
    f = tmp_path / "r.egsinp"
    write_isource1_input(f)
    txt = f.read_text()
    assert REC13 in txt
    f.write_text(txt.replace(REC13, rec))
    p, err = parse(f)
    t = transport_options(p)
    assert (t["ireject"], t["esave"]) == (ireject, esave)
    if note:
        assert note in err


def run(tp, n, ireject):
    xb = np.linspace(-4, 4, 17)
    zb = np.linspace(0, 10, 21)
    med = np.zeros(16 * 16 * 20, np.int32)
    kz = np.arange(med.size) // 256
    med[(kz >= 6) & (kz < 10)] = 2   # bone at 3-5 cm
    src = make_source(0, 1.25, x0=-3.5, x1=3.5, y0=-3.5, y1=3.5, z=-1.0)
    tp.ireject = ireject
    tp.esave = 2.0
    sim = Simulation(Tables(BLOB521), xb, xb.copy(), zb, med, None, src, tp, nstreams=16384)
    sim.run(n)
    t, c = sim.tallies(), sim.counters()
    sim.close()
    g, keep = _lib.make_geometry(xb, xb.copy(), zb, med, None)
    return t, c, (g, keep, src)


@pytest.mark.gpu
def test_range_rejection_keeps_the_dose_and_cuts_electron_steps():
    """1.25 MeV photons on 0.5 cm voxels of water with a bone slab, ECUT
    0.521, ESAVE_GLOBAL 2 MeV: dose against the oracle, electron steps
    against the same engine without rejection."""
    tp = _lib.default_transport(ecut=0.521, pcut=0.01)
    n = 800000
    b, cb, (g, keep, src) = run(tp, n, 1)
    _, c0, _ = run(_lib.default_transport(ecut=0.521, pcut=0.01), n, 0)
    o = oracle_lib.OracleSimulation(BLOB521, g, src[0], _lib.default_transport(ecut=0.521,
                                                                               pcut=0.01),
                                    nstreams=256, nthreads=8)
    o.run(n)
    a = o.tallies()
    o0 = oracle_lib.OracleSimulation(BLOB521, g, src[0], _lib.default_transport(ecut=0.521,
                                                                                pcut=0.01),
                                     nstreams=256, nthreads=8)
    o0.seed(4711, 20000)
    o0.run(n)
    print("range rejection: electron steps %d -> %d" % (c0[4], cb[4]))
    check_parity(a, b, cb, label="range rejection", total_rtol=0.01, null=o0.tallies())
    assert cb[4] < 0.7 * c0[4]


@pytest.mark.gpu
def test_range_rejection_combinations():
    xb = np.linspace(-2, 2, 5)
    med = np.zeros(64, np.int32)
    tp = _lib.default_transport(ecut=0.521, pcut=0.01, ireject=1, esave=2.0, howfarless=1)
    with pytest.raises(_lib.EngineError, match="HOWFARLESS"):
        Simulation(Tables(BLOB521), xb, xb.copy(), xb + 2, med, None,
                   make_source(0, 1.0, z=-1.0), tp)
    tp = _lib.default_transport(ecut=0.521, pcut=0.01, ireject=1, esave=2.0)
    with pytest.raises(_lib.EngineError, match="plain history"):
        Simulation(Tables(BLOB521), xb, xb.copy(), xb + 2, med, None,
                   make_source(0, 1.0, z=-1.0), tp, schedule="event")

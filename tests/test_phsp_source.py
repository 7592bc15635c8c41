"""Phase-space sources: EGSnrc MODE0/MODE2 files (the records
EGS_PhspSource reads, egs++/sources/egs_phsp_source/egs_phsp_source.cpp
:79-470) in Python and in the host application, and phase-space particles
driving the transport (egsgpu_set_phsp) against the oracle's parallel beams
with the same distribution."""
import os
import subprocess

import numpy as np
import pytest

from egsnrc_amd import _lib
from egsnrc_amd.engine import Simulation, Tables, make_phsp_source, make_source
from egsnrc_amd.phantom_io import RM, read_egsphsp, write_egsphsp
from tests import oracle_lib, workloads
from tests.test_host_app import APP, BLOB700

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BLOB521 = os.path.join(ROOT, "egsnrc_amd", "data", "521icru.egstab")


def test_phsp_round_trip_and_byte_order(tmp_path):
    n = 7
    e = np.array([0.5, 1.0, 2.0, 3.0, 0.2, 4.0, 6.0])
    q = np.array([0, -1, 1, 0, 0, -1, 0])
    x = np.linspace(-1, 1, n)
    wt = np.array([1, -1, 2, 1, 1, 0.5, 1.0])
    f = tmp_path / "a.egsphsp"
    write_egsphsp(f, e, x, -x, 0.1 * x, 0.0 * x, wt, q, pinc=3.0,
                  new_history=np.array([1, 0, 0, 1, 0, 1, 0], bool))
    p = read_egsphsp(f)
    assert (p["n"], p["nphot"], p["pinc"]) == (7, 4, 3.0)
    assert np.allclose(p["e"], e, atol=1e-6) and list(p["q"]) == list(q)
    assert list(p["new_history"]) == [True, False, False, True, False, True, False]
    # charged particles are stored with their total energy
    raw = np.frombuffer(open(f, "rb").read()[28:], dtype="<f4").reshape(n, 7)
    assert np.allclose(np.abs(raw[:, 1]), e + np.where(q != 0, RM, 0), atol=1e-5)
    # the same file in the other byte order reads the same
    b = bytearray(open(f, "rb").read())
    head = np.frombuffer(bytes(b[5:25]), "<i4").byteswap().tobytes()
    body = np.frombuffer(bytes(b[28:]), "<i4").byteswap().tobytes()
    g = tmp_path / "b.egsphsp"
    g.write_bytes(bytes(b[:5]) + head + bytes(b[25:28]) + body)
    p2 = read_egsphsp(g)
    assert p2["n"] == 7 and np.array_equal(p2["e"], p["e"]) and np.array_equal(p2["q"], p["q"])


INPUT = """
:start geometry definition:
    :start geometry:
        library = egs_ndgeometry
        type = EGS_XYZGeometry
        name = phantom
        x-planes = -2 -1 0 1 2
        y-planes = -2 -1 0 1 2
        z-planes = 0 1 2 3
        :start media input:
            media = H2O700ICRU
        :stop media input:
    :stop geometry:
    simulation geometry = phantom
:stop geometry definition:
:start source definition:
    :start source:
        library = egs_phsp_source
        name = the_phsp
        phase space file = %s
        particle type = %s
    :stop source:
    :start source:
        library = egs_transformed_source
        name = placed
        source name = the_phsp
        :start transformation:
            translation = 0.5 -0.25 -1
        :stop transformation:
    :stop source:
    simulation source = placed
:stop source definition:
:start run control:
    ncase = 1000
:stop run control:
"""


def test_app_reads_a_phase_space_source(tmp_path):
    n = 12
    q = np.array([0, -1, 1, 0] * 3)
    write_egsphsp(tmp_path / "p.egsphsp", np.full(n, 1.0), np.zeros(n), np.zeros(n),
                  np.zeros(n), np.zeros(n), np.ones(n), q, pinc=6.0)
    f = tmp_path / "p.egsinp"
    f.write_text(INPUT % ("p.egsphsp", "photons"))
    r = subprocess.run([APP, "-i", str(f), "--parse-only"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    out = {l.partition(":")[0]: l.partition(":")[2].split() for l in r.stdout.splitlines()}
    # particles, electrons, photons, positrons, rejected (not photons), incident
    assert out["phsp"] == ["12", "3", "6", "3", "6", "6"]
    assert [float(v) for v in out["phsp plane"]] == [0.5, -0.25, -1.0]
    assert out["source"][:2] == ["kind", "3"]


def beam_particles(n, rng, half=2.0):
    return (rng.uniform(-half, half, n), rng.uniform(-half, half, n), np.zeros(n), np.zeros(n))


def oracle_run(xb, yb, zb, med, src, tp, n, blob):
    g, keep = _lib.make_geometry(xb, yb, zb, med, None)
    o = oracle_lib.OracleSimulation(blob, g, src[0], tp, nstreams=256, nthreads=8)
    o.run(n)
    return o.tallies()


@pytest.mark.gpu
def test_phsp_weights_and_rejections_match_the_oracle():
    """1.25 MeV photons of a 4 x 4 cm parallel beam given as phase-space
    particles, alternately of weight 2 and of weight 0 (a rejected
    particle): the same expected dose as the oracle's parallel beam."""
    n = 400000
    rng = np.random.default_rng(5)
    x, y, u, v = beam_particles(n, rng)
    wt = np.where(np.arange(n) % 2 == 0, 2.0, 0.0)
    xb = np.linspace(-4, 4, 9)
    zb = np.linspace(0, 10, 11)
    med = np.zeros(8 * 8 * 10, np.int32)
    tp = _lib.default_transport(ecut=0.7, pcut=0.01)
    sim = Simulation(Tables(BLOB700), xb, xb.copy(), zb, med, None, make_phsp_source(z=-1.0), tp,
                     nstreams=8192, schedule="history")
    sim.set_phsp(x, y, u, v, np.full(n, 1.25), wt, np.zeros(n))
    sim.run(n)
    b = sim.tallies()
    a = oracle_run(xb, xb.copy(), zb, med, make_source(0, 1.25, x0=-2, x1=2, y0=-2, y1=2,
                                                         z=-1.0), tp, n, BLOB700)
    chi2, out2, nv = workloads.compare_doses(a, b)
    print("phsp weights: chi2/dof %.3f  >2sigma %.3f  voxels %d" % (chi2, out2, nv))
    assert nv > 60
    assert chi2 < max(1.3, 1 + 3 * np.sqrt(2.0 / nv))
    assert out2 < max(0.09, 0.0455 + 3 * np.sqrt(0.0455 * 0.9545 / nv))


@pytest.mark.gpu
def test_mixed_phsp_on_the_event_schedule_matches_the_oracle():
    """A phase space alternating 2 MeV photons and 6 MeV electrons (entering
    the photon bank and handed to the charged-particle bank): the dose per
    history is half the oracle's photon-beam dose plus half its
    electron-beam dose."""
    n = 200000
    rng = np.random.default_rng(6)
    x, y, u, v = beam_particles(n, rng)
    q = np.where(np.arange(n) % 2 == 0, 0, -1)
    e = np.where(q == 0, 2.0, 6.0)
    xb = np.linspace(-4, 4, 17)
    zb = np.linspace(0, 10, 21)
    med = np.zeros(16 * 16 * 20, np.int32)
    kz = np.arange(16 * 16 * 20) // 256
    med[(kz == 6) | (kz == 7)] = 2   # a bone slab at 3-4 cm
    tp = _lib.default_transport(ecut=0.521, pcut=0.01)
    sim = Simulation(Tables(BLOB521), xb, xb.copy(), zb, med, None, make_phsp_source(z=-1.0), tp,
                     nstreams=8192, schedule="event")
    sim.set_phsp(x, y, u, v, e, np.ones(n), q)
    sim.run(n)
    b = sim.tallies()
    c = sim.counters()
    assert c[5] == 0 and c[6] == 0
    half = n // 2
    tabs = []
    for charge, energy in ((0, 2.0), (-1, 6.0)):
        src = make_source(charge, energy, x0=-2, x1=2, y0=-2, y1=2, z=-1.0)
        tabs.append(oracle_run(xb, xb.copy(), zb, med, src, tp, half, BLOB521))
    mb = b[0] / b[2]
    vb = np.maximum(b[1] / b[2] - mb * mb, 0) / (b[2] - 1)
    ma = 0.5 * sum(t[0] / t[2] for t in tabs)
    va = 0.25 * sum(np.maximum(t[1] / t[2] - (t[0] / t[2]) ** 2, 0) / (t[2] - 1) for t in tabs)
    keep = (np.maximum(ma, mb) > 0.1 * max(ma.max(), mb.max())) & (va + vb > 0)
    z = (ma[keep] - mb[keep]) / np.sqrt(va[keep] + vb[keep])
    chi2, out2, nv = float(np.mean(z * z)), float(np.mean(np.abs(z) > 2)), int(keep.sum())
    print("mixed phsp: chi2/dof %.3f  >2sigma %.3f  voxels %d" % (chi2, out2, nv))
    assert nv > 300
    assert chi2 < max(1.3, 1 + 3 * np.sqrt(2.0 / nv))
    assert out2 < max(0.09, 0.0455 + 3 * np.sqrt(0.0455 * 0.9545 / nv))

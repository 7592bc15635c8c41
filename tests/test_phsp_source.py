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
from egsnrc_amd import workloads
from tests import oracle_lib
from tests.parity import check_parity, check_parity_means
from tests.test_host_app import APP, BLOB700, read_3ddose

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


def oracle_run(xb, yb, zb, med, src, tp, n, blob, seeds=None):
    g, keep = _lib.make_geometry(xb, yb, zb, med, None)
    o = oracle_lib.OracleSimulation(blob, g, src[0], tp, nstreams=256, nthreads=8)
    if seeds:
        o.seed(*seeds)
    o.run(n)
    return o.tallies()


@pytest.mark.gpu
def test_phsp_weights_and_rejections_match_the_oracle():
    """1.25 MeV photons of a 6 x 6 cm parallel beam given as phase-space
    particles, alternately of weight 2 and of weight 0 (a rejected
    particle): the same expected dose as the oracle's parallel beam."""
    n = 800000
    rng = np.random.default_rng(5)
    x, y, u, v = beam_particles(n, rng, half=3.0)
    wt = np.where(np.arange(n) % 2 == 0, 2.0, 0.0)
    xb = np.linspace(-4, 4, 17)
    zb = np.linspace(0, 10, 21)
    med = np.zeros(16 * 16 * 20, np.int32)
    tp = _lib.default_transport(ecut=0.7, pcut=0.01)
    sim = Simulation(Tables(BLOB700), xb, xb.copy(), zb, med, None, make_phsp_source(z=-1.0), tp,
                     nstreams=8192, schedule="history")
    sim.set_phsp(x, y, u, v, np.full(n, 1.25), wt, np.zeros(n))
    sim.run(n)
    b = sim.tallies()
    beam = make_source(0, 1.25, x0=-3, x1=3, y0=-3, y1=3, z=-1.0)
    a = oracle_run(xb, xb.copy(), zb, med, beam, tp, n, BLOB700)
    a0 = oracle_run(xb, xb.copy(), zb, med, beam, tp, n, BLOB700, seeds=(4711, 20000))
    check_parity(a, b, sim.counters(), label="phsp weights", null=a0)


@pytest.mark.gpu
def test_mixed_phsp_on_the_event_schedule_matches_the_oracle():
    """A phase space alternating 2 MeV photons and 6 MeV electrons (entering
    the photon bank and handed to the charged-particle bank): the dose per
    history is half the oracle's photon-beam dose plus half its
    electron-beam dose."""
    n = 400000
    rng = np.random.default_rng(6)
    x, y, u, v = beam_particles(n, rng, half=5.5)
    q = np.where(np.arange(n) % 2 == 0, 0, -1)
    e = np.where(q == 0, 2.0, 6.0)
    xb = np.linspace(-6, 6, 25)
    zb = np.linspace(0, 10, 21)
    med = np.zeros(24 * 24 * 20, np.int32)
    kz = np.arange(24 * 24 * 20) // (24 * 24)
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
        src = make_source(charge, energy, x0=-5.5, x1=5.5, y0=-5.5, y1=5.5, z=-1.0)
        tabs.append(oracle_run(xb, xb.copy(), zb, med, src, tp, half, BLOB521))
    mb = b[0] / b[2]
    vb = np.maximum(b[1] / b[2] - mb * mb, 0) / (b[2] - 1)
    ma = 0.5 * sum(t[0] / t[2] for t in tabs)
    va = 0.25 * sum(np.maximum(t[1] / t[2] - (t[0] / t[2]) ** 2, 0) / (t[2] - 1) for t in tabs)
    check_parity_means(ma, np.sqrt(va), mb, np.sqrt(vb), "mixed phsp")


def write_iaea(base, e, x, y, z, u, v, wt, q, new_history=None, byteorder="<",
               store_z=False, orig=6.0):
    """An IAEA phase space as the IAEA library writes one
    (iaea_record.cpp:86-135, header blocks iaea_header.cpp:720-810): the
    sign of w in the type, a new history as a negative energy; x, y, u, v,
    weight stored, z stored or a RECORD_CONSTANT."""
    n = len(e)
    nh = np.zeros(n, bool) if new_history is None else new_history
    typ = np.choose(np.asarray(q) + 1, [2, 1, 3]).astype(np.int8)
    w = np.sqrt(np.maximum(0, 1 - np.asarray(u) ** 2 - np.asarray(v) ** 2))
    sign = np.where(np.asarray(wt) < 0, -1, 1)
    typ = (typ * sign).astype(np.int8)
    cols = [np.where(nh, -np.asarray(e), e), x, y] + ([z] if store_z else []) + \
        [u, v, np.abs(wt)]
    fl = np.stack([np.asarray(c, np.float32) for c in cols], 1).astype(byteorder + "f4")
    rec = np.zeros(n, dtype=[("t", "i1"), ("f", byteorder + "f4", (fl.shape[1],))])
    rec["t"] = typ
    rec["f"] = fl
    open(str(base) + ".IAEAphsp", "wb").write(rec.tobytes())
    zc = "" if store_z else "   %g     // Constant Z\n" % z[0]
    hdr = ("$IAEA_INDEX:\n   0\n\n$TITLE:\n   test\n\n$FILE_TYPE:\n   0\n\n"
           "$CHECKSUM:\n   %d\n\n$RECORD_CONTENTS:\n    1     // X is stored ?\n"
           "    1     // Y is stored ?\n    %d     // Z is stored ?\n    1     // U is stored ?\n"
           "    1     // V is stored ?\n    1     // W is stored ?\n"
           "    1     // Weight is stored ?\n    0     // Extra floats stored ?\n"
           "    0     // Extra longs stored ?\n\n$RECORD_CONSTANT:\n%s\n"
           "$RECORD_LENGTH:\n   %d\n\n$BYTE_ORDER:\n   %s\n\n$ORIG_HISTORIES:\n   %g\n\n"
           "$PARTICLES:\n   %d\n\n" % (rec.nbytes, 1 if store_z else 0, zc,
                                      rec.dtype.itemsize, "1234" if byteorder == "<" else "4321",
                                      orig, n))
    open(str(base) + ".IAEAheader", "w").write(hdr)


IAEA_INPUT = INPUT.replace("library = egs_phsp_source", "library = iaea_phsp_source") \
    .replace("phase space file = %s", "iaea phase space file = %s")


def parse_phsp(path):
    r = subprocess.run([APP, "-i", str(path), "--parse-only"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    return {l.partition(":")[0]: l.partition(":")[2].split() for l in r.stdout.splitlines()}


@pytest.mark.parametrize("byteorder,store_z", [("<", False), (">", True)])
def test_iaea_file_reads_as_the_same_particles(tmp_path, byteorder, store_z):
    """The same particles as an EGSnrc MODE0 file and as an IAEA file (z
    constant or stored, either byte order) read to the same arrays."""
    n = 40
    rng = np.random.default_rng(8)
    e = rng.uniform(0.1, 6, n)
    x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    u, v = rng.uniform(-0.3, 0.3, n), rng.uniform(-0.3, 0.3, n)
    wt = rng.uniform(0.5, 2, n) * np.where(np.arange(n) % 7 == 3, -1, 1)
    q = rng.choice([-1, 0, 1], n)
    write_egsphsp(tmp_path / "p.egsphsp", e, x, y, u, v, wt, q, pinc=6.0)
    write_iaea(tmp_path / "p", e, x, y, np.full(n, 1.0), u, v, wt, q, byteorder=byteorder,
               store_z=store_z)
    f1 = tmp_path / "a.egsinp"
    f1.write_text(INPUT % ("p.egsphsp", "all"))
    f2 = tmp_path / "b.egsinp"
    f2.write_text(IAEA_INPUT % ("p", "all"))
    a, b = parse_phsp(f1), parse_phsp(f2)
    assert a["phsp"] == b["phsp"] and a["phsp sums"][1:] == b["phsp sums"][1:]
    # MODE0 stores a charged particle's total energy in FP32: its kinetic
    # energy comes back rounded there
    assert float(a["phsp sums"][0]) == pytest.approx(float(b["phsp sums"][0]), rel=1e-7)
    # the IAEA particles sit on their own plane z = 1, translated by -1
    assert [float(t) for t in b["phsp plane"]] == [0.5, -0.25, 0.0]


def test_iaea_header_mismatch_is_an_error(tmp_path):
    n = 5
    write_iaea(tmp_path / "p", np.ones(n), np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n),
               np.zeros(n), np.ones(n), np.zeros(n, int))
    h = (tmp_path / "p.IAEAheader").read_text().replace("$RECORD_LENGTH:\n   25",
                                                       "$RECORD_LENGTH:\n   29")
    (tmp_path / "p.IAEAheader").write_text(h)
    f = tmp_path / "b.egsinp"
    f.write_text(IAEA_INPUT % ("p", "all"))
    r = subprocess.run([APP, "-i", str(f), "--parse-only"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode != 0 and "RECORD_LENGTH" in r.stderr


@pytest.mark.gpu
def test_iaea_and_egsphsp_sources_give_the_same_dose(tmp_path):
    """One photon list as MODE0 and as IAEA (z = 0 plane): the same dose
    within statistics (the event schedule hands particles to streams in an
    order that varies from run to run, so the runs are not bit-identical)."""
    n = 3000
    rng = np.random.default_rng(9)
    e = rng.uniform(0.5, 3, n)
    x, y = rng.uniform(-1.5, 1.5, n), rng.uniform(-1.5, 1.5, n)
    q = np.zeros(n, int)
    write_egsphsp(tmp_path / "p.egsphsp", e, x, y, np.zeros(n), np.zeros(n), np.ones(n), q,
                  pinc=float(n))
    write_iaea(tmp_path / "p", e, x, y, np.zeros(n), np.zeros(n), np.zeros(n), np.ones(n), q,
               orig=float(n))
    outs = []
    for name, txt in (("a", INPUT % ("p.egsphsp", "all")), ("b", IAEA_INPUT % ("p", "all"))):
        f = tmp_path / (name + ".egsinp")
        f.write_text(txt.replace("ncase = 1000", "ncase = 20000"))
        r = subprocess.run([APP, "-i", str(f), "-p", BLOB700], capture_output=True, text=True,
                           cwd=tmp_path, timeout=300)
        assert r.returncode == 0, r.stderr
        outs.append(read_3ddose(str(tmp_path / (name + ".3ddose"))))
    (_, _, da, ea), (_, _, db, eb) = outs
    m = np.maximum(da, db) > 0.05 * da.max()
    z = (da[m] - db[m]) / np.sqrt((da[m] * ea[m]) ** 2 + (db[m] * eb[m]) ** 2)
    print("iaea vs egsphsp: chi2/dof %.3f over %d voxels" % (np.mean(z * z), m.sum()))
    assert m.sum() >= 20 and np.mean(z * z) < 1 + 3 * np.sqrt(2.0 / m.sum())

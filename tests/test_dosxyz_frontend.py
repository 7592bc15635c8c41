"""The DOSXYZnrc front end of the host application: DOSXYZnrc input
records (user_codes/dosxyznrc/dosxyznrc.mortran:150-330), sources isource 0
and 1 (srcxyznrc.mortran:104-146, 2442-2539, 3416-3448, 3551-3571,
4255-4392), .spectrum files with the inverse-CDF grid (:1811-1899) and the
.3ddose normalised per incident fluence (dosxyznrc.mortran:2997-3130).

The parse checks run on the CPU (--parse-only); the GPU tests run the
reference's own example input and compare the doses per incident fluence
with the oracle run with the equivalent egs++ parallel beam."""
import os
import subprocess

import numpy as np
import pytest

from egsnrc_amd import _lib
from egsnrc_amd.engine import make_source
from tests import oracle_lib
from tests.parity import check_parity_means
from tests.test_host_app import APP, BLOB700, read_3ddose

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXAMPLE = os.path.join(ROOT, "tests", "golden", "dosxyznrc_air_bone_interface_example.egsinp")
MEV_J = 1.602e-10   # the MeV -> J (times 1e3 g -> kg) factor DOSXYZnrc uses


def parse(path):
    r = subprocess.run([APP, "-i", str(path), "--parse-only"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        k, _, v = line.partition(":")
        out[k.strip()] = v.split()
    return out, r.stderr


def test_parse_reference_example():
    p, err = parse(EXAMPLE)
    assert p["format"] == ["dosxyznrc"]
    assert p["grid"] == ["21", "21", "20"]
    assert [float(v) for v in p["x"]] == [-5.25, 5.25]
    assert [float(v) for v in p["z"]] == [0.0, 10.0]
    assert p["media"] == ["H2O700ICRU", "ICRPBONE700ICRU", "AIR700ICRU"]
    # bone: i, j 8..14 and k 5..10 (7*7*6); air: k 4 only (7*7*1)
    assert p["voxels per medium"] == [str(21 * 21 * 20 - 294 - 49), "294", "49"]
    assert p["source"][:4] == ["kind", "1", "charge", "-1"]
    src = dict(zip(p["source"][::2], p["source"][1::2]))
    assert float(src["energy"]) == 20.0
    assert float(src["area"]) == pytest.approx(2.5 * 4.5)
    assert [float(v) for v in p["source rectangle"][:4]] == [-1.25, 1.25, -2.25, 2.25]
    assert p["ncase"] == ["500000"]
    # INSEED1 = -97 is out of range and takes the default 1802
    assert p["seeds"] == ["1802", "33"]
    assert float(p["ecut"][0]) == 0.7 and float(p["pcut"][0]) == 0.01
    # PRESTA-I boundary crossing with mscati's default skin depth, PRESTA-II
    # steps
    t = transport_options(p)
    assert (t["exact_bca"], t["skin"], t["step_algorithm"]) == ("0", "0", "0")


def write_isource1_input(path, spectrum=None, xcol=3.0, ycol=2.0, ncase=100000, seeds=(33, 97)):
    """8 x 8 x 10 cm water phantom of 1 cm voxels with a bone block
    (i, j 3..6, k 4..6), a beam along +z (theta = 180) through the
    isocentre (0, 0, 5) with collimator widths xcol x ycol."""
    lines = ["isource 1 test", "2", "H2O700ICRU", "ICRPBONE700ICRU",
             "0.7, 0.01, 0, 0, 0", "-1, -1, -1, 0",
             "-4.0", "1.0, 8", "-4.0", "1.0, 8", "0.0", "1.0, 10",
             "3, 6, 3, 6, 4, 6, 2, 0.0",
             "0, 0, 0, 0, 0, 0, 0, 0",
             "0, 0, 0, 0, 0, 0, 0, 0",
             "0, 0, 0, 0, 0, 0, 0, 0",
             "0, 1, 0.0, 0.0, 5.0, 180.0, 0.0, %g, %g, 0.0" % (xcol, ycol),
             "%d, 0, 0, 0, 0, 0, 0, 0" % (1 if spectrum else 0),
             str(spectrum) if spectrum else "1.25",
             "%d, 0, 500.0, %d, %d, 100.0, 0, 0, 0, 0, , 0, 0, 0, 0, 0, 0" % (ncase, *seeds),
             " :Start MC Transport Parameter:",
             " Global ECUT= 0.7", " Global PCUT= 0.01", " ESTEPE= 0.25",
             " Spin effects= Off", " Bound Compton scattering= Off",
             " Rayleigh scattering= Off", " Atomic relaxations= Off",
             " Photoelectron angular sampling= Off",
             " :Stop MC Transport Parameter:"]
    path.write_text("\n".join(lines) + "\n")


def test_parse_isource1_geometry_of_the_beam(tmp_path):
    f = tmp_path / "s1.egsinp"
    write_isource1_input(f)
    p, _ = parse(f)
    src = dict(zip(p["source"][::2], p["source"][1::2]))
    assert src["kind"] == "2" and src["charge"] == "0"
    assert float(src["area"]) == pytest.approx(6.0)
    u = [float(v) for v in p["direction"]]
    assert u[2] == pytest.approx(1.0) and abs(u[0]) < 1e-6 and abs(u[1]) < 1e-6
    # farthest corner of the grid from the isocentre (0, 0, 5): (4, 4, 5)
    assert float(p["dsource"][0]) == pytest.approx(-np.sqrt(16 + 16 + 25 + 9 + 4))
    assert p["voxels per medium"] == [str(640 - 48), "48"]


def test_spectrum_inverse_cdf_grid(tmp_path):
    """Two bins with cumulative probability exactly 0.25 at the first edge:
    grid point K (probability K/1000) belongs to the first bin while
    K/1000 <= 0.25, i.e. K = 1..250 (srcxyznrc.mortran:1889-1899)."""
    spec = tmp_path / "two.spectrum"
    spec.write_text("two bins\n2, 0.5, 0\n1.0, 0.25\n2.0, 0.75\n")
    f = tmp_path / "s1.egsinp"
    write_isource1_input(f, spectrum=spec)
    p, _ = parse(f)
    c = p["cdfinv"]
    assert c[0] == "1000"
    n_first, lo1, w1, lo2, w2 = int(c[1]), float(c[2]), float(c[3]), float(c[4]), float(c[5])
    assert n_first == 250
    assert (lo1, w1) == (0.5, 0.5) and (lo2, w2) == (1.0, 1.0)


def per_fluence_from_3ddose(path, rho, area):
    """Energy per incident particle (MeV) and its sigma from a DOSXYZnrc
    .3ddose (dose per fluence in Gy cm^2, relative errors)."""
    dims, b, d, e = read_3ddose(path)
    vol = np.einsum("k,j,i->kji", np.diff(b[2]), np.diff(b[1]), np.diff(b[0])).reshape(-1)
    E = d * rho * vol / (MEV_J * area)
    return E, e * E


def oracle_energy(g, src, tp, n, seeds=None):
    o = oracle_lib.OracleSimulation(BLOB700, g, src[0], tp, nstreams=256, nthreads=8)
    if seeds:
        o.seed(*seeds)
    o.run(n)
    s, s2, nn = o.tallies()
    m = s / nn
    v = np.maximum(s2 / nn - m * m, 0) / (nn - 1)
    return m, np.sqrt(v)


def compare(ma, sa, mb, sb, label, min_voxels, null=None):
    """Parity of the front end's run with the oracle's (the shared bar; the
    phantoms of these inputs are small, so min_voxels is theirs)."""
    check_parity_means(mb, sb, ma, sa, label, min_voxels=min_voxels, null=null)


def transport(**kw):
    return _lib.default_transport(ecut=0.7, pcut=0.01, rayleigh=0, bound_compton=0,
                                  photoelectron_angle=0, atomic_relaxation=0, **kw)


@pytest.mark.gpu
def test_reference_example_runs_and_matches_the_oracle(tmp_path):
    """The DOSXYZnrc example (20 MeV electrons, isource 0, water with an air
    layer and a bone block) run unchanged but for NCASE, against the oracle
    with the same phantom and the equivalent parallel beam."""
    n = 400000
    txt = open(EXAMPLE).read().replace("500000, 0, 500.0", "%d, 0, 500.0" % n)
    f = tmp_path / "air_bone.egsinp"
    f.write_text(txt)
    r = subprocess.run([APP, "-i", str(f), "-p", BLOB700], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    # the phantom of the example: 21 x 21 x 20 voxels of 0.5 cm, bone
    # (1.85 g/cm3) at i, j 8..14, k 5..10 and air at i, j 8..14, k 4
    xb = np.linspace(-5.25, 5.25, 22)
    zb = np.linspace(0.0, 10.0, 21)
    med = np.zeros((20, 21, 21), np.int32)
    med[4:10, 7:14, 7:14] = 1
    med[3, 7:14, 7:14] = 2
    rho = np.choose(med, [1.0, 1.85, 0.0012048]).reshape(-1)
    med = med.reshape(-1)
    E, sE = per_fluence_from_3ddose(tmp_path / "air_bone.3ddose", rho, 2.5 * 4.5)
    g, keep = _lib.make_geometry(xb, xb.copy(), zb, med, None)
    src = make_source(-1, 20.0, x0=-1.25, x1=1.25, y0=-2.25, y1=2.25, z=-1.0)
    # "Brems angular sampling= Simple", "Boundary crossing algorithm=
    # PRESTA-I" and "Skin depth for BCA= 0" in the example
    tp = transport(brems_angle=0, exact_bca=0, skin_depth=0.0)
    m, s = oracle_energy(g, src, tp, n)
    null = oracle_energy(g, src, tp, n, seeds=(4711, 20000))
    compare(E, sE, m, s, "dosxyznrc example, isource 0", 200, null=null)


@pytest.mark.gpu
def test_isource1_spectrum_and_misses_match_the_oracle(tmp_path):
    """isource 1 photon beam along +z with an energy spectrum (enflag 1) and
    a collimator twice the phantom face, so that three quarters of the beam
    misses: the dose per incident fluence (misses counted, as nmissm) equals
    the oracle's for a beam exactly covering the face."""
    spec = tmp_path / "falling.spectrum"
    # counts per bin, cumulative probabilities multiples of 1/1000
    spec.write_text("falling\n4, 0.25, 0\n0.5, 0.4\n1.0, 0.3\n2.0, 0.2\n4.0, 0.1\n")
    f = tmp_path / "s1.egsinp"
    n = 400000
    write_isource1_input(f, spectrum=spec, xcol=16.0, ycol=16.0, ncase=n)
    r = subprocess.run([APP, "-i", str(f), "-p", BLOB700], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    nmiss = [int(l.split(":")[1]) for l in r.stdout.splitlines() if "missed" in l]
    assert nmiss and 2.5 * n < nmiss[0] < 3.5 * n
    med = np.zeros((10, 8, 8), np.int32)
    med[3:6, 2:6, 2:6] = 1
    rho = np.choose(med, [1.0, 1.85]).reshape(-1)
    E, sE = per_fluence_from_3ddose(tmp_path / "s1.3ddose", rho, 16.0 * 16.0)
    xb = np.linspace(-4, 4, 9)
    g, keep = _lib.make_geometry(xb, xb.copy(), np.linspace(0, 10, 11), med.reshape(-1), None)
    edges = np.array([0.25, 0.5, 1.0, 2.0, 4.0])
    per_mev = np.array([0.4, 0.3, 0.2, 0.1]) / np.diff(edges)
    src = make_source(0, 0.0, spectrum=(edges, per_mev), x0=-4, x1=4, y0=-4, y1=4, z=-1.0)
    no = 100000
    m, s = oracle_energy(g, src, transport(), no)
    # both per unit fluence: the oracle's beam covers the 8 x 8 cm face
    compare(E, sE, m * 64.0 / (16.0 * 16.0), s * 64.0 / (16.0 * 16.0), "isource 1 + spectrum",
            100)


def test_parse_howfarless_only_for_a_homogeneous_phantom(tmp_path):
    """ihowfarless (record 13, field 16) switches HOWFARLESS on for a
    homogeneous phantom and is reported and ignored otherwise."""
    f = tmp_path / "h.egsinp"
    write_isource1_input(f)
    txt = f.read_text().replace("100000, 0, 500.0, 33, 97, 100.0, 0, 0, 0, 0, , 0, 0, 0, 0, 0, 0",
                                "100000, 0, 500.0, 33, 97, 100.0, 0, 0, 0, 0, , 0, 0, 0, 0, 1, 0")
    f.write_text(txt)
    p, err = parse(f)
    assert transport_options(p)["howfarless"] == "0" and "not homogeneous" in err
    g = tmp_path / "w.egsinp"
    g.write_text(txt.replace("3, 6, 3, 6, 4, 6, 2, 0.0", "3, 6, 3, 6, 4, 6, 1, 0.0"))
    p, err = parse(g)
    assert transport_options(p)["howfarless"] == "1"


def transport_options(p):
    t = p["transport"]
    return dict(zip(t[::2], t[1::2]))


def test_parse_photon_splitting(tmp_path):
    """n_split (record 13, field 15) is handed to the transport; 0 and 1
    leave photons unsplit."""
    f = tmp_path / "s.egsinp"
    write_isource1_input(f)
    assert transport_options(parse(f)[0])["n_split"] == "0"
    txt = f.read_text().replace("100000, 0, 500.0, 33, 97, 100.0, 0, 0, 0, 0, , 0, 0, 0, 0, 0, 0",
                                "100000, 0, 500.0, 33, 97, 100.0, 0, 0, 0, 0, , 0, 0, 0, 40, 0, 0")
    assert txt != f.read_text()
    f.write_text(txt)
    assert transport_options(parse(f)[0])["n_split"] == "40"


def write_isource3_input(path, field, ssd, ncase=100000):
    write_isource1_input(path, ncase=ncase)
    txt = path.read_text().replace("0, 1, 0.0, 0.0, 5.0, 180.0, 0.0, 3, 2, 0.0",
                                   "0, 3, %g, %g, %g, %g, %g" % (*field, ssd))
    assert "0, 3, " in txt
    path.write_text(txt)


def test_parse_isource3_clamps_the_field(tmp_path):
    """isource 3 (srcxyznrc.mortran:1100-1109, 2550-2586): the field is
    clamped to the front face, beamarea its area."""
    f = tmp_path / "s3.egsinp"
    write_isource3_input(f, (-6.0, 2.0, -1.5, 1.5), 100.0)
    p, _ = parse(f)
    src = dict(zip(p["source"][::2], p["source"][1::2]))
    assert src["kind"] == "4"
    assert [float(v) for v in p["source rectangle"][:4]] == [-4.0, 2.0, -1.5, 1.5]
    assert float(src["area"]) == pytest.approx(18.0)
    assert float(p["dsource"][0]) == 100.0


@pytest.mark.gpu
@pytest.mark.parametrize("field,area", [((-2.0, 2.0, -2.0, 2.0), 16.0), ((0.5, 0.5, 0.5, 0.5), 0)])
def test_isource3_far_point_source_matches_a_parallel_beam(tmp_path, field, area):
    """A frontal point source 10^5 cm away is a parallel beam: over a 4 x 4
    cm field it gives the oracle's 4 x 4 cm parallel beam, and with a field
    of zero width a pencil beam at (0.5, 0.5)."""
    f = tmp_path / "s3.egsinp"
    n = 200000
    write_isource3_input(f, field, 1.0e5, ncase=n)
    r = subprocess.run([APP, "-i", str(f), "-p", BLOB700], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    med = np.zeros((10, 8, 8), np.int32)
    med[3:6, 2:6, 2:6] = 1
    rho = np.choose(med, [1.0, 1.85]).reshape(-1)
    E, sE = per_fluence_from_3ddose(tmp_path / "s3.3ddose", rho, area if area else 1.0)
    xb = np.linspace(-4, 4, 9)
    g, keep = _lib.make_geometry(xb, xb.copy(), np.linspace(0, 10, 11), med.reshape(-1), None)
    if area:
        src = make_source(0, 1.25, x0=-2, x1=2, y0=-2, y1=2, z=-1.0)
    else:
        src = make_source(0, 1.25, shape="point", x0=0.5, y0=0.5, z=-1.0)
    m, s = oracle_energy(g, src, transport(), n)
    compare(E, sE, m, s, "isource 3 field %s" % (field,), 40 if area else 8)

"""Reference file formats: .egsphant phantoms, CT ramps, .3ddose files and
the .3ddose comparator (egsnrc_amd/phantom_io.py, host app reader)."""
import os
import subprocess

import numpy as np
import pytest

from egsnrc_amd import phantom_io as P
from egsnrc_amd.engine import write_3ddose

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
PHANT = os.path.join(GOLD, "ndgeom_egsphant.egsphant.gz")   # egs++ geometry example
RAMP = os.path.join(GOLD, "ndgeom_egsphant.ramp")
APP = os.path.join(ROOT, "egsnrc_amd", "egs_gpu_app")
BLOB521 = os.path.join(ROOT, "egsnrc_amd", "data", "521icru.egstab")


def test_read_reference_example_phantom():
    """The egs++ example phantom (geometry/examples/ndgeom_egsphant.*): 53^3
    voxels, 0.2 cm inside +-5.1 cm with one 9.9 cm voxel either side; water
    density (0.998) for z < 0 and air (0.0012) above.  egs++ ignores the
    medium characters and takes media from densities + CT ramp
    (egs_nd_geometry.cpp:688-697, 740-764); in this example the characters
    label the two halves the other way round from the ramp."""
    d = P.read_egsphant(PHANT)
    assert d["media"] == ["MEDIUM_1", "MEDIUM_2"]
    assert d["shape"] == (53, 53, 53)
    for b in (d["xb"], d["yb"], d["zb"]):
        assert b[0] == -15 and b[-1] == 15 and b[1] == -5.1 and b[-2] == 5.1
        np.testing.assert_allclose(np.diff(b[1:-1]), 0.2, atol=1e-9)
    ramp = P.read_ramp(RAMP)
    assert [r[0] for r in ramp] == ["MEDIUM_1", "MEDIUM_2"]
    med, rel = P.apply_ramp(d["density"], ramp)
    assert np.all(med >= 0)
    np.testing.assert_array_equal(med, 1 - d["medium"])
    z = 0.5 * (d["zb"][1:] + d["zb"][:-1])
    rho = d["density"].reshape(53, 53, 53)
    assert np.all(rho[z < 0] == 0.998) and np.all(rho[z > 0] == 0.0012)
    # densities 0.0012 (rho_default 0.12) and 0.998 (rho_default 0.998)
    np.testing.assert_allclose(np.unique(rel), [0.01, 1.0])


def test_egsphant_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    xb = np.array([-1.0, 0.0, 0.5, 2.0]); yb = np.array([0.0, 1.0, 3.0]); zb = np.array([0.0, 0.25])
    med = rng.integers(0, 3, size=3 * 2 * 1)
    rho = np.round(rng.uniform(0.1, 2.0, size=6), 4)
    f = str(tmp_path / "p.egsphant")
    P.write_egsphant(f, ["A", "B", "C"], xb, yb, zb, med, rho)
    d = P.read_egsphant(f)
    np.testing.assert_array_equal(d["medium"], med)
    np.testing.assert_allclose(d["density"], rho)
    np.testing.assert_allclose(d["xb"], xb)
    assert d["shape"] == (3, 2, 1)


def test_ramp_ranges_and_vacuum():
    ramp = [("AIR", 0.0, 0.1, 0.0012), ("WATER", 0.9, 1.1, 1.0)]
    med, rel = P.apply_ramp([0.0012, 0.05, 1.0, 1.05, 0.5, 1.1], ramp)
    np.testing.assert_array_equal(med, [0, 0, 1, 1, -1, -1])   # rho_max is exclusive
    np.testing.assert_allclose(rel[:4], [1.0, 0.05 / 0.0012, 1.0, 1.05])


def test_3ddose_round_trip_and_comparator(tmp_path):
    rng = np.random.default_rng(2)
    xb = np.linspace(-2, 2, 5); yb = np.linspace(-1, 1, 3); zb = np.linspace(0, 3, 4)
    n = 4 * 2 * 3
    mass = np.full(n, 0.5)
    edep = rng.uniform(1.0, 2.0, n)
    dedep = 0.02 * edep
    fa = str(tmp_path / "a.3ddose")
    write_3ddose(fa, xb, yb, zb, edep, dedep, mass)
    ga = P.read_3ddose(fa)
    np.testing.assert_allclose(ga[3], edep * 1.602e-10 / mass, rtol=1e-5)   # %g text
    np.testing.assert_allclose(ga[4], 0.02, rtol=1e-5)
    same = P.compare_3ddose(fa, fa)
    assert same["chi2_per_dof"] == 0 and same["voxels"] == n
    # an independent run with the stated uncertainty: z-scores ~ N(0,1); a
    # comparator that used relative instead of absolute errors, or dropped
    # one run's uncertainty, would not give chi2/dof near 1
    m = 20000
    xb2 = np.linspace(0, m, m + 1)
    d1 = np.full(m, 100.0); s1 = 2.0
    d2 = d1 + rng.normal(0, np.sqrt(2) * s1, m)
    f1, f2 = str(tmp_path / "1.3ddose"), str(tmp_path / "2.3ddose")
    write_3ddose(f1, xb2, [0, 1], [0, 1], d1, np.full(m, s1), np.full(m, 1.602e-10))
    write_3ddose(f2, xb2, [0, 1], [0, 1], d2, np.full(m, s1) * d2 / d1, np.full(m, 1.602e-10))
    c = P.compare_3ddose(f1, f2)
    assert 0.95 < c["chi2_per_dof"] < 1.05
    assert 0.035 < c["frac_beyond_2sigma"] < 0.057
    with pytest.raises(ValueError):
        P.compare_3ddose(fa, f1)


def egsphant_input(phant, ramp, ncase):
    return """
:start geometry definition:
    :start geometry:
        name = phantom
        library = egs_ndgeometry
        type = EGS_XYZGeometry
        egsphant file = %s
        ct ramp = %s
    :stop geometry:
    simulation geometry = phantom
:stop geometry definition:
:start source definition:
    :start source:
        library = egs_parallel_beam
        name = beam
        charge = 0
        direction = 0 0 1
        :start shape:
            type = rectangle
            rectangle = -2 -2 2 2
            :start transformation:
                translation = 0 0 -20
            :stop transformation:
        :stop shape:
        :start spectrum:
            type = monoenergetic
            energy = 2.0
        :stop spectrum:
        :stop source:
    simulation source = beam
:stop source definition:
:start run control:
    ncase = %d
    nbatch = 2
:stop run control:
:start MC transport parameter:
    Global ECUT = 0.521
    Global PCUT = 0.01
    Spin effects = Off
:stop MC transport parameter:
:start ausgab object definition:
    :start ausgab object:
        name = dose
        library = egs_dose_scoring
        output dose file = yes
    :stop ausgab object:
:stop ausgab object definition:
""" % (phant, ramp, ncase)


def test_app_reads_egsphant_with_ramp_errors(tmp_path):
    """The ramp of the example names media the table blob does not hold."""
    f = tmp_path / "ph.egsinp"
    f.write_text(egsphant_input(PHANT, RAMP, 1000))
    r = subprocess.run([APP, "-i", str(f), "-p", BLOB521], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode != 0 and "MEDIUM_1" in r.stderr


@pytest.mark.gpu
def test_app_runs_egsphant_with_ramp(tmp_path):
    ramp = tmp_path / "water_air.ramp"
    ramp.write_text("AIR521ICRU 0.0 0.5 0.0012\nH2O521ICRU 0.8 1.1 0.998\n")
    f = tmp_path / "ph.egsinp"
    f.write_text(egsphant_input(PHANT, str(ramp), 100000))
    r = subprocess.run([APP, "-i", str(f), "-p", BLOB521], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    xb, yb, zb, d, e = P.read_3ddose(str(tmp_path / "ph.3ddose"))
    assert len(d) == 53 ** 3
    d3 = d.reshape(53, 53, 53)
    ph = P.read_egsphant(PHANT)
    water = ph["density"].reshape(53, 53, 53) > 0.5
    # the beam enters the water half (z < 0) of the central column and
    # deposits there; beside the beam (|x| > 7 cm) the water stays nearly dark
    col = d3[:, 26, 26][water[:, 26, 26]]
    assert np.all(col[1:] > 0)
    assert d3[:, 26, 0][water[:, 26, 0]].max() < 0.1 * col.max()

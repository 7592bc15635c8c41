"""Polar angles of the pair particles ($SET-PAIR-ANGLE, egsnrc.macros
:2070-2163; "Pair angular sampling" Off / Simple / KM / Uniform / Blend,
egs_advanced_application.cpp:546-551), from the oracle's PAIR with the
photon along +z.  Each option's distribution is checked through its
cumulative distribution (the probability-integral transform must be
uniform):

- Simple: the leading term of the angular distribution, whose cosine c in
  the particle's rest frame is isotropic, so z = (E w - p) / (E - p w) is
  uniform on [-1, 1] (E, p the particle's total energy and momentum);
- Uniform: w uniform on [-1, 1];
- Blend: per particle the leading term with probability E / (E + 10 MeV),
  else w = 1 - 2 sqrt(r), i.e. F(w) = 1 - ((1 - w) / 2)^2;
- KM: the leading term below 4.14 MeV; above it only the bounds and the
  scale of the angles (theta < pi; the median of theta E / m between 0.2
  and 3) are checked, not the shape of the 3D-2003 distribution.
The engine and the host app accept exactly these five options.  GPU: dose
parity of the engine with the oracle under 20 MeV photons with KM and Blend
angles."""
import ctypes as C
import os

import numpy as np
import pytest
from scipy import stats

from egsnrc_amd import _lib
from egsnrc_amd.engine import Tables, make_source
from tests import oracle_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BLOB = os.path.join(ROOT, "egsnrc_amd", "data", "521icru.egstab")
RM = 0.5109989461
WATER = 0


def sample(energy, option, n=100000):
    out = oracle_lib.sample(BLOB, 14, WATER, energy, n, param=float(option))
    return out[:, :2].reshape(-1), out[:, 2:].reshape(-1)   # energies, cosines


def leading_term_cdf(w, e):
    p = np.sqrt(np.maximum((e - RM) * (e + RM), 0))
    z = (e * w - p) / (e - p * w)
    return (np.clip(z, -1, 1) + 1) / 2


def test_simple_leading_term():
    for energy in (1.5, 6.0, 30.0):
        e, w = sample(energy, 1)
        assert stats.kstest(leading_term_cdf(w, e), "uniform").pvalue > 1e-3, energy


def test_uniform():
    e, w = sample(10.0, 3)
    assert stats.kstest((w + 1) / 2, "uniform").pvalue > 1e-3


def test_blend():
    for energy in (5.0, 40.0):
        e, w = sample(energy, 4)
        pl = e / (e + 10)
        u = pl * leading_term_cdf(w, e) + (1 - pl) * (1 - ((1 - w) / 2) ** 2)
        assert stats.kstest(u, "uniform").pvalue > 1e-3, energy


def test_km():
    # below 4.14 MeV: the leading term
    e, w = sample(3.0, 2)
    assert stats.kstest(leading_term_cdf(w, e), "uniform").pvalue > 1e-3
    for energy in (6.0, 20.0, 60.0):
        e, w = sample(energy, 2, n=50000)
        th = np.arccos(np.clip(w, -1, 1))
        assert np.all(th < np.pi)
        med = np.median(th * e / RM)
        assert 0.2 < med < 3.0, (energy, med)
    # and not the leading term's distribution (the two approach each other
    # at higher energies)
    e, w = sample(6.0, 2)
    assert stats.kstest(leading_term_cdf(w, e), "uniform").pvalue < 1e-6


def test_off_keeps_the_photon_direction():
    e, w = sample(10.0, 0, n=1000)
    assert np.all(w == 1.0)


def test_options_outside_the_five_are_refused(tmp_path):
    L = _lib.lib()
    xb = np.array([0.0, 1.0])
    tab = Tables(BLOB)
    g, keep = _lib.make_geometry(xb, xb.copy(), xb.copy(), np.zeros(1, np.int32), None)
    src = make_source(0, 10.0)
    for bad in (dict(pair_angle=5), dict(pair_angle=-1), dict(brems_angle=2)):
        tp = _lib.default_transport(**bad)
        h = C.c_void_p()
        rc = L.egsgpu_create(tab.h, C.byref(g), C.byref(src[0]), C.byref(tp), C.byref(h))
        # the device check comes first where there is no GPU
        assert rc != 0
    from tests.test_dosxyz_frontend import EXAMPLE, parse, transport_options
    txt = open(EXAMPLE).read()
    for name, k in (("Off", 0), ("Simple", 1), ("KM", 2), ("Uniform", 3), ("Blend", 4)):
        f = tmp_path / ("pa_%s.egsinp" % name)
        f.write_text(txt.replace("Pair angular sampling= Simple", "Pair angular sampling= " + name))
        p, _ = parse(f)
        assert transport_options(p)["pair_angle"] == str(k)


def pair_case(n=24, dx=0.5, energy=20.0, half=3.0, pair_angle=2):
    """20 MeV photons (6 x 6 cm^2) on a 12 cm water cube in 5 mm voxels with
    the given pair angular sampling (ECUT 0.521 MeV total, PCUT 0.01 MeV)."""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    med = np.zeros(n ** 3, dtype=np.int32)
    src = make_source(0, energy, x0=-half, x1=half, y0=-half, y1=half, z=-1.0)
    tp = _lib.default_transport(kerma=0, ecut=0.521, pcut=0.01, pair_angle=pair_angle)
    return xb, xb.copy(), zb, med, None, src, tp


HIST = 400000
_ORACLE = {}


def _oracle(pair_angle, seeds=None):
    key = (pair_angle, seeds)
    if key not in _ORACLE:
        xb, yb, zb, med, rhor, src, tp = pair_case(pair_angle=pair_angle)
        g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
        o = oracle_lib.OracleSimulation(BLOB, g, src[0], tp, nstreams=256,
                                        nthreads=max(1, min(os.cpu_count() or 1, 16)))
        if seeds:
            o.seed(*seeds)
        o.run(HIST)
        _ORACLE[key] = o.tallies()
    return _ORACLE[key]


@pytest.mark.gpu
@pytest.mark.parametrize("pair_angle", [2, 4])
def test_pair_angle_dose_parity(engine, pair_angle):
    from egsnrc_amd.engine import Simulation
    from tests.parity import check_parity
    xb, yb, zb, med, rhor, src, tp = pair_case(pair_angle=pair_angle)
    sim = Simulation(Tables(BLOB), xb, yb, zb, med, rhor, src, tp, nstreams=8192)
    sim.run(HIST)
    b = sim.tallies()
    assert b[2] == HIST
    c = sim.counters()
    assert c[5] == 0 and c[6] == 0
    check_parity(_oracle(pair_angle), b, c, label="pair angle %d" % pair_angle,
                 null=_oracle(pair_angle, (4242, 2718)))

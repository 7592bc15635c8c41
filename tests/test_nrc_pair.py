"""NRC pair cross sections (pair_nrc = 1; init_nrc_pair,
egsnrc.mortran:7256-7352, and the NRC branch of PAIR, :4596-4620).

CPU: the oracle's positron energy shares at tabulated photon energies
against the densities of data/pair_nrc1.data for hydrogen and oxygen
(fixture tests/golden/nrc_pair_water.json), mixed here with the weights
n_Z Z^2 and integrated piecewise linearly over the 64 intervals of the
fraction grid; the Coulomb asymmetry (positrons take more than half near
threshold); the option needs a blob with the tables.  GPU: dose parity of
the engine with the oracle under a 15 MeV photon beam."""
import ctypes as C
import json
import os

import numpy as np
import pytest
from scipy import stats

from egsnrc_amd import _lib
from egsnrc_amd.engine import Simulation, Tables, make_source
from tests import oracle_lib
from tests.parity import check_parity
from tests.test_oracle_pinning import read_blob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BLOB = os.path.join(ROOT, "egsnrc_amd", "data", "521icru_nist.egstab")
BLOB_BH = os.path.join(ROOT, "egsnrc_amd", "data", "521icru.egstab")
FIX = os.path.join(ROOT, "tests", "golden", "nrc_pair_water.json")
RM = 0.5109989461
WATER = 0


def fraction_grid(nb):
    """init_nrc_pair's positron energy-fraction points: quadratic from 0 to
    1/2 and mirrored to 1."""
    nbb = nb // 2
    dd = np.sqrt(0.5) / nbb
    lo = (dd * np.arange(nbb + 1)) ** 2
    return np.concatenate([lo, 1 - lo[:-1][::-1]])


def test_positron_energy_shares_follow_the_data():
    fx = json.load(open(FIX))
    b = read_blob(BLOB)
    z, pz = b["m0.z"], b["m0.pz"]
    x = fraction_grid(fx["nb"])
    assert np.allclose(x, b["nrcp_xdata"], atol=1e-15)
    dle = np.log((fx["emax"] - 2) / (fx["emin"] - 2)) / (fx["ne"] - 1)
    for ie in (1, 5, 20, 40, 60):
        k = 2 + (fx["emin"] - 2) * np.exp((ie - 1) * dle)
        f = sum(pj * zj * zj * np.array(fx["f_by_Z"][str(int(zj + 0.5))][str(ie)])
                for zj, pj in zip(z, pz))
        out = oracle_lib.sample(BLOB, 4, WATER, k * RM, 300000)
        # the positron's kinetic-energy share (o[2], o[3] are the charges)
        ep = np.where(out[:, 2] > 0, out[:, 0], out[:, 1])
        xx = (ep / RM - 1) / (k - 2)
        assert np.all((xx >= -1e-9) & (xx <= 1 + 1e-9))
        # expected counts per grid interval: the linear density's area
        mass = 0.5 * (f[1:] + f[:-1]) * np.diff(x)
        h, _ = np.histogram(xx, x)
        expct = mass / mass.sum() * len(xx)
        ok = expct > 20
        chi2 = np.sum((h[ok] - expct[ok]) ** 2 / expct[ok])
        pval = stats.chi2.sf(chi2, ok.sum() - 1)
        print("k %.4f: chi2 %.1f / %d, <x+> %.4f" % (k, chi2, ok.sum() - 1, xx.mean()))
        assert pval > 1e-3, (ie, chi2)
        assert np.allclose(out[:, 0] + out[:, 1], k * RM, atol=1e-9)
        if ie <= 20:
            # near threshold the nucleus pushes the positron: it takes more
            # than half the kinetic energy on average (Bethe-Heitler is
            # symmetric)
            assert xx.mean() > 0.5 + 5 * xx.std() / np.sqrt(len(xx))


def test_pair_option_needs_the_tables():
    L = _lib.lib()
    xb = np.array([0.0, 1.0])
    tab = Tables(BLOB_BH)
    g, keep = _lib.make_geometry(xb, xb.copy(), xb.copy(), np.zeros(1, np.int32), None)
    src = make_source(0, 10.0)
    tp = _lib.default_transport(pair_xsec=1)
    h = C.c_void_p()
    assert L.egsgpu_create(tab.h, C.byref(g), C.byref(src[0]), C.byref(tp), C.byref(h)) == -1
    buf = C.create_string_buffer(512)
    L.egsgpu_last_error(buf, 512)
    assert b"pair" in buf.value


def nrc_case(n=30, dx=0.5, energy=15.0, half=4.0):
    """15 MeV photons (8 x 8 cm^2) on water with a bone slab at 2-4 cm
    depth in 5 mm voxels, NIST brems and NRC pair cross sections (ECUT
    0.521 MeV total, PCUT 0.01 MeV)."""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    zc = 0.5 * (zb[1:] + zb[:-1])
    mz = np.where((zc > 2.0) & (zc < 4.0), 1, 0)
    med = np.broadcast_to(mz[:, None, None], (n, n, n)).reshape(-1).astype(np.int32)
    src = make_source(0, energy, x0=-half, x1=half, y0=-half, y1=half, z=-1.0)
    tp = _lib.default_transport(kerma=0, ecut=0.521, pcut=0.01, brems_xsec=1, pair_xsec=1)
    return xb, xb.copy(), zb, med, None, src, tp


HIST = 400000
NULL_SEEDS = (4242, 2718)
_ORACLE = {}


def _oracle(seeds=None):
    if seeds not in _ORACLE:
        xb, yb, zb, med, rhor, src, tp = nrc_case()
        g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
        o = oracle_lib.OracleSimulation(BLOB, g, src[0], tp, nstreams=256,
                                        nthreads=max(1, min(os.cpu_count() or 1, 16)))
        if seeds:
            o.seed(*seeds)
        o.run(HIST)
        _ORACLE[seeds] = o.tallies()
    return _ORACLE[seeds]


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["event", "history"])
def test_nrc_pair_dose_parity(engine, schedule):
    xb, yb, zb, med, rhor, src, tp = nrc_case()
    sim = Simulation(Tables(BLOB), xb, yb, zb, med, rhor, src, tp, nstreams=8192,
                     schedule=schedule)
    sim.run(HIST)
    b = sim.tallies()
    assert b[2] == HIST
    c = sim.counters()
    assert c[5] == 0 and c[6] == 0
    check_parity(_oracle(), b, c, label="nrc pair/" + schedule, null=_oracle(NULL_SEEDS))

"""EADL atomic relaxation (egs_eadl_relax, egsnrc.mortran:6020-6216, called
by PHOTO and COMPT): the oracle's cascade is pinned by energy conservation,
by the K-shell fluorescence yields and X-ray line energies of lead and
tantalum (Krause's yields, the X-ray data booklet's line energies), and by
the frequencies of the K-shell radiative transitions against the EADL
probabilities as relax.data holds them (tests/golden/relax_k_shells.json,
written by tests/golden/make_relax_fixture.py); the engine is then held to
the oracle on a photon beam that makes lead fluoresce into water."""
import json
import os

import numpy as np
import pytest
from scipy import stats

from egsnrc_amd import _lib
from egsnrc_amd.engine import Simulation, Tables, make_source
from tests import oracle_lib
from tests.blob import read_blob
from tests.parity import check_parity

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TUTOR = os.path.join(ROOT, "egsnrc_amd", "data", "tutor.egstab")   # TA, PB, H2O
FIX = os.path.join(ROOT, "tests", "golden", "relax_k_shells.json")
RELAX = 11   # oracle_sample kind: one cascade from a vacancy


def cascades(Z, shell, n, seed=17):
    """n cascades from a vacancy in `shell` of element Z: first product
    energy (negative: electron), counts, emitted kinetic energy, local."""
    return oracle_lib.sample(TUTOR, RELAX, 1, float(Z), n, param=float(shell), seed=seed)


def shell_be(Z):
    b = read_blob(TUTOR)
    z0 = int(b["relax_zfirst"][Z])
    return b["shell_be"][z0:z0 + int(b["relax_nshell"][Z])]


@pytest.mark.parametrize("Z,shell", [(82, 1), (82, 2), (82, 4), (73, 1), (73, 3)])
def test_cascade_conserves_energy(Z, shell):
    """Every cascade gives the vacancy's binding energy to the products on
    the stack and to local absorption, to rounding."""
    out = cascades(Z, shell, 20000)
    be = shell_be(Z)[shell - 1]
    np.testing.assert_allclose(out[:, 2] + out[:, 3], be, rtol=1e-12, atol=0)
    assert out[:, 1].max() < 1e6 * 8   # the vacancy stack never holds more than 8


@pytest.mark.parametrize("Z,omega_krause,kev", [(82, 0.963, 50.0), (73, 0.957, 40.0)])
def test_k_fluorescence_yield(Z, omega_krause, kev):
    """The share of K vacancies filled by a radiative transition (the first
    product is a K x-ray) equals the sum of the radiative EADL
    probabilities within 3 binomial sigma, and Krause's fluorescence yield
    (J. Phys. Chem. Ref. Data 8 (1979) 307) within 1 %."""
    n = 400000
    out = cascades(Z, 1, n)
    frac = float(np.mean(out[:, 0] > kev / 1000))
    k = json.load(open(FIX))[str(Z)]["K"]
    st, pr = np.array(k["states"]), np.array(k["probs"])
    omega = pr[st < 64].sum() / pr.sum()
    print("Z %d: K radiative share %.4f, EADL %.4f, Krause %.3f" % (Z, frac, omega, omega_krause))
    assert abs(frac - omega) < 3 * np.sqrt(omega * (1 - omega) / n)
    assert abs(omega - omega_krause) < 0.01


def test_lead_k_lines():
    """Lead's K x-rays come out at the booklet energies (K-alpha1 74.969,
    K-alpha2 72.805, K-beta1 84.938 keV; within 50 eV, the difference
    between the XCOM edge and the EADL binding energies) and with the EADL
    branching: chi^2 of the counts of every radiative K line (final shell)
    against the relax.data probabilities."""
    n = 400000
    out = cascades(82, 1, n, seed=23)
    e = out[:, 0][out[:, 0] > 0.05] * 1000   # keV
    vals, cnt = np.unique(np.round(e, 3), return_counts=True)
    top = vals[np.argsort(-cnt)[:3]]
    for ref in (74.969, 72.805, 84.938):
        assert np.min(np.abs(top - ref)) < 0.05, (top, ref)
    k = json.load(open(FIX))["82"]["K"]
    st, pr = np.array(k["states"]), np.array(k["probs"]) / np.sum(k["probs"])
    be = shell_be(82) * 1000
    rad = st < 64
    exp_e = be[0] - be[st[rad]]          # state counted from 0: final shell st + 1
    obs = np.array([np.sum(np.abs(e - x) < 1e-3) for x in exp_e])
    expct = pr[rad] * n
    keep = expct > 20
    chi2 = float(np.sum((obs[keep] - expct[keep]) ** 2 / expct[keep]))
    dof = int(keep.sum()) - 1
    p = stats.chi2.sf(chi2, dof)
    print("K lines: chi2 %.1f / %d dof, p = %.3g" % (chi2, dof, p))
    assert p > 1e-3
    # every K x-ray is one of the radiative lines
    assert obs.sum() == e.size


def lead_case(n=32, dx=0.1):
    """200 keV photons (2.4 x 2.4 cm^2) on water with a 2 mm lead slab at
    0.5-0.7 cm depth: the lead absorbs most photons in its K shell and its
    K x-rays (75-85 keV) carry energy into the water on either side."""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    zc = 0.5 * (zb[1:] + zb[:-1])
    mz = np.where((zc > 0.5) & (zc < 0.7), 1, 2)        # PB, H2O of the blob
    med = np.broadcast_to(mz[:, None, None], (n, n, n)).reshape(-1).astype(np.int32)
    src = make_source(0, 0.2, x0=-1.2, x1=1.2, y0=-1.2, y1=1.2, z=-1.0)
    tp = _lib.default_transport(ecut=0.7, pcut=0.01, kerma=0)
    return xb, xb.copy(), zb, med, None, src, tp


LEAD_HIST = 10_000_000
_RUNS = {}


def lead_oracle(seeds=None):
    if seeds not in _RUNS:
        xb, yb, zb, med, rhor, src, tp = lead_case()
        g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
        o = oracle_lib.OracleSimulation(TUTOR, g, src[0], tp, nstreams=256,
                                        nthreads=max(1, min(os.cpu_count() or 1, 16)))
        if seeds:
            o.seed(*seeds)
        o.run(LEAD_HIST)
        _RUNS[seeds] = o.tallies()
    return _RUNS[seeds]


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["history", "event"])
def test_lead_fluorescence_parity(engine, schedule):
    """The engine's cascade against the oracle's: voxel parity over the
    water and the lead, and the total deposit per history within 0.3 %."""
    xb, yb, zb, med, rhor, src, tp = lead_case()
    sim = Simulation(Tables(TUTOR), xb, yb, zb, med, rhor, src, tp, schedule=schedule)
    sim.run(LEAD_HIST)
    b = sim.tallies()
    check_parity(lead_oracle(), b, sim.counters(), label="lead fluorescence/" + schedule,
                 frac=1e-3, total_rtol=3e-3, null=lead_oracle((4711, 20000)))

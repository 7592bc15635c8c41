"""NIST and NRC bremsstrahlung cross sections (ibr_nist = 1 and 2;
init_nist_brems, egsnrc.mortran:6893-7226, and the NIST branch of BREMS,
:404-422).

CPU: the table builder's total brems cross section of water and the
oracle's photon-energy sampling are checked against the scaled cross
sections of hydrogen and oxygen -- Seltzer and Berger's (fixture
tests/golden/nist_brems_water.json, from data/nist_brems.data) and
Tessier's (tests/golden/nrc_brems_water.json, from data/nrc_brems.data) --
integrated and interpolated independently here with scipy's natural cubic
splines; the engine refuses a blob whose brems tables differ from the
option.  GPU: dose parity of the engine with the oracle under a 10 MeV
electron beam, with either set."""
import ctypes as C
import json
import os

import numpy as np
import pytest
from scipy import integrate, stats
from scipy.interpolate import CubicSpline

from egsnrc_amd import _lib
from egsnrc_amd.engine import Simulation, Tables, make_source
from tests import oracle_lib
from tests.parity import check_parity
from tests.test_oracle_pinning import read_blob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BLOB = os.path.join(ROOT, "egsnrc_amd", "data", "521icru_nist.egstab")
BLOB_NRC = os.path.join(ROOT, "egsnrc_amd", "data", "521icru_nrcbrems.egstab")
BLOB_BH = os.path.join(ROOT, "egsnrc_amd", "data", "521icru.egstab")
FIX = os.path.join(ROOT, "tests", "golden", "nist_brems_water.json")
FIX_NRC = os.path.join(ROOT, "tests", "golden", "nrc_brems_water.json")
# brems_xsec -> (blob, fixture)
SETS = {1: (BLOB, FIX), 2: (BLOB_NRC, FIX_NRC)}
RM = 0.5109989461
WATER = 0
AMU = 1660.5655   # mb/atom -> cm^2/g for A in g/mol


def water(ibr=1):
    """Scaled cross section of water per gram on the data grid: sum over
    the elements of n_Z Z^2 chi_Z(T, k/T) / (sum n_Z A_Z amu)."""
    blob, fix = SETS[ibr]
    fx = json.load(open(fix))
    b = read_blob(blob)
    assert int(b["ibr_nist"][0]) == ibr
    z, pz, wa = b["m0.z"], b["m0.pz"], b["m0.wa"]
    T, x = np.array(fx["T"]), np.array(fx["x"])
    suma = np.sum(pz * wa) * AMU
    cs = np.zeros((len(x), len(T)))
    for zj, pj in zip(z, pz):
        cs += pj * zj * zj / suma * np.array(fx["cs_by_Z"][str(int(zj + 0.1))])
    return T, x, cs, b


def spectrum_spline(x, cs_col):
    """log cs over xi = log(1 - k/T + 1e-6), natural cubic spline (xi made
    increasing for scipy)."""
    xi = np.log(1 - x + 1e-6)
    return CubicSpline(xi[::-1], np.log(cs_col)[::-1], bc_type="natural")


@pytest.mark.parametrize("ibr", [1, 2])
def test_total_brems_cross_section_of_water(ibr):
    """sigma_brems(T) = rho / beta^2 int_AP^T chi(T, k/T) dk / k: the blob's
    esig * ebr1 at the electron-table nodes against the integral of the
    data (quadrature here; data energies joined by a natural spline in
    log T, as the builder joins them)."""
    T, x, cs, b = water(ibr)
    ap, rho = b["m0.ap"][0], b["m0.rho"][0]
    sel = T > ap * 1.5
    sig_data = []
    for n in np.nonzero(sel)[0]:
        s = spectrum_spline(x, cs[:, n])
        aux = np.log(T[n] / ap)
        f = lambda u: np.exp(s(np.log(1 - ap / T[n] * np.exp(u * aux) + 1e-6)))
        sig_data.append(integrate.quad(f, 0, 1, limit=400, epsabs=0, epsrel=1e-10)[0])
    inter = CubicSpline(np.log(T[sel]), sig_data, bc_type="natural")
    eke0, eke1, meke = b["m0.eke0"][0], b["m0.eke1"][0], int(b["m0.meke"][0])
    checked = 0
    for i in range(2, meke + 1):
        el = (i - eke0) / eke1
        e = np.exp(el)
        if not (0.03 < e < 50.0):
            continue
        # node i is the upper end of the linear piece i-1 (1-based)
        sig = (b["m0.esig1"][i - 2] * el + b["m0.esig0"][i - 2]) * \
              (b["m0.ebr11"][i - 2] * el + b["m0.ebr10"][i - 2])
        beta2 = e * (e + 2 * RM) / (e + RM) ** 2
        want = inter(el) * np.log(e / ap) / beta2 * rho
        assert abs(sig - want) < 2e-4 * want, (e, sig, want)
        checked += 1
    assert checked > 100


@pytest.mark.parametrize("ibr", [1, 2])
def test_nist_photon_energies_follow_the_data(ibr):
    """BREMS with the NIST tables: br = log(k/AP) / log(T/AP) is drawn from
    the alias table of the grid energy T (50 linear bins refined where
    they deviate most); its density must follow chi(T, k/T) of the data
    (linear in log T between data energies, natural spline in
    log(1 - k/T))."""
    T, x, cs, b = water(ibr)
    ap = b["m0.ap"][0]
    lemin, dlei = b["m0.nb_lemin"][0], b["m0.nb_dlei"][0]
    for j in (5, 30, 60, 95):
        e = np.exp(lemin + (j - 1) / dlei)
        n = np.searchsorted(T, e, side="right") - 1
        n = min(max(n, 0), len(T) - 2)
        ple = (np.log(e) - np.log(T[n])) / (np.log(T[n + 1]) - np.log(T[n]))
        col = (1 - ple) * cs[:, n] + ple * cs[:, n + 1]
        s = spectrum_spline(x, col)
        aux = np.log(e / ap)
        pdf = lambda u: np.exp(s(np.log(1 - ap / e * np.exp(u * aux) + 1e-6)))
        out = oracle_lib.sample(SETS[ibr][0], 7, WATER, e, 400000)
        k = out[:, 0]
        assert np.all(k >= ap * (1 - 1e-9)) and np.all(k <= e * (1 + 1e-9))
        br = np.log(k / ap) / aux
        edges = np.linspace(0, 1, 41)
        p = np.array([integrate.quad(pdf, a0, a1, limit=200)[0]
                      for a0, a1 in zip(edges[:-1], edges[1:])])
        h, _ = np.histogram(br, edges)
        expct = p / p.sum() * len(br)
        chi2 = np.sum((h - expct) ** 2 / expct)
        pval = stats.chi2.sf(chi2, len(h) - 1)
        print("T %.4g MeV chi2 %.1f / %d" % (e, chi2, len(h) - 1))
        assert pval > 1e-3, (e, chi2)
        # energy conservation of the pair of particles
        assert np.allclose(out[:, 0] + out[:, 1], e + RM, atol=1e-9)


def test_engine_refuses_mismatched_brems_tables():
    """The option and the blob must agree (checked before any device
    work): NIST asked of Bethe-Heitler tables, and the reverse."""
    L = _lib.lib()
    xb = np.array([0.0, 1.0])
    for path, nist in ((BLOB_BH, 1), (BLOB, 0), (BLOB_NRC, 1), (BLOB, 2), (BLOB_BH, 2)):
        tab = Tables(path)
        g, keep = _lib.make_geometry(xb, xb.copy(), xb.copy(), np.zeros(1, np.int32), None)
        src = make_source(-1, 1.0)
        tp = _lib.default_transport(brems_xsec=nist)
        h = C.c_void_p()
        rc = L.egsgpu_create(tab.h, C.byref(g), C.byref(src[0]), C.byref(tp), C.byref(h))
        assert rc == -1
        buf = C.create_string_buffer(512)
        L.egsgpu_last_error(buf, 512)
        assert b"brems" in buf.value


def nist_case(n=30, dx=0.4, energy=10.0, half=4.0, ibr=1):
    """10 MeV electrons (8 x 8 cm^2) on water with a bone slab at 1-2 cm
    depth in 4 mm voxels, NIST (ibr 1) or NRC (ibr 2) brems cross sections
    (ECUT 0.521 MeV total, PCUT 0.01 MeV)."""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    zc = 0.5 * (zb[1:] + zb[:-1])
    mz = np.where((zc > 1.0) & (zc < 2.0), 1, 0)
    med = np.broadcast_to(mz[:, None, None], (n, n, n)).reshape(-1).astype(np.int32)
    src = make_source(-1, energy, x0=-half, x1=half, y0=-half, y1=half, z=-1.0)
    tp = _lib.default_transport(kerma=0, ecut=0.521, pcut=0.01, brems_xsec=ibr)
    return xb, xb.copy(), zb, med, None, src, tp


def test_oracle_nist_run_deposits_what_bh_deposits():
    """NIST's brems cross section of water is 5-11 % above Bethe-Heitler's
    in the number of photons above AP (the tables above), but the two agree
    on the radiated energy to a few per cent, so a 10 MeV beam leaves the
    same energy in the phantom (within 0.3 %) with either; a sampler that
    drew photon energies from the wrong table would not."""
    xb, yb, zb, med, rhor, src, tp = nist_case()
    res = {}
    for name, path, m, t in (("nist", BLOB, med, tp),
                             ("bh", BLOB_BH, np.where(med == 1, 2, 0).astype(np.int32),
                              _lib.default_transport(kerma=0, ecut=0.521, pcut=0.01))):
        g, keep = _lib.make_geometry(xb, yb, zb, m, rhor)
        o = oracle_lib.OracleSimulation(path, g, src[0], t, nstreams=64, nthreads=8)
        o.run(40000)
        s, s2, n = o.tallies()
        res[name] = s.sum() / n
    print(res)
    assert 8.0 < res["nist"] < 10.0
    assert abs(res["nist"] - res["bh"]) < 3e-3 * res["bh"]


HIST = 200000
NULL_SEEDS = (4242, 2718)
_ORACLE = {}


def _oracle(seeds=None, ibr=1):
    key = (seeds, ibr)
    if key not in _ORACLE:
        xb, yb, zb, med, rhor, src, tp = nist_case(ibr=ibr)
        g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
        o = oracle_lib.OracleSimulation(SETS[ibr][0], g, src[0], tp, nstreams=256,
                                        nthreads=max(1, min(os.cpu_count() or 1, 16)))
        if seeds:
            o.seed(*seeds)
        o.run(HIST)
        _ORACLE[key] = o.tallies()
    return _ORACLE[key]


def test_oracle_nrc_run_deposits_what_nist_deposits():
    """The NRC and NIST scaled cross sections differ by a few per cent at
    most below 10 MeV, so a 10 MeV beam leaves the same energy in the
    phantom with either set (within 0.3 %)."""
    res = {}
    for ibr in (1, 2):
        xb, yb, zb, med, rhor, src, tp = nist_case(ibr=ibr)
        g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
        o = oracle_lib.OracleSimulation(SETS[ibr][0], g, src[0], tp, nstreams=64, nthreads=8)
        o.run(40000)
        s, s2, n = o.tallies()
        res[ibr] = s.sum() / n
    print(res)
    assert abs(res[2] - res[1]) < 3e-3 * res[1]


@pytest.mark.gpu
@pytest.mark.parametrize("schedule,ibr", [("event", 1), ("history", 1), ("event", 2)])
def test_nist_brems_dose_parity(engine, schedule, ibr):
    xb, yb, zb, med, rhor, src, tp = nist_case(ibr=ibr)
    sim = Simulation(Tables(SETS[ibr][0]), xb, yb, zb, med, rhor, src, tp, nstreams=8192,
                     schedule=schedule)
    sim.run(HIST)
    b = sim.tallies()
    assert b[2] == HIST
    c = sim.counters()
    assert c[5] == 0 and c[6] == 0
    check_parity(_oracle(ibr=ibr), b, c, label="brems %d/%s" % (ibr, schedule),
                 null=_oracle(NULL_SEEDS, ibr))


def test_host_app_reads_the_brems_cross_section_option(tmp_path):
    """`Brems cross sections = NIST` (NRC) in the transport block selects
    the NIST (NRC) tables."""
    from tests.test_dosxyz_frontend import EXAMPLE, parse, transport_options
    txt = open(EXAMPLE).read()
    f = tmp_path / "nist.egsinp"
    f.write_text(txt.replace("Brems cross sections= BH", "Brems cross sections= NIST"))
    p, _ = parse(f)
    assert transport_options(p)["brems_xsec"] == "1"
    f.write_text(txt.replace("Brems cross sections= BH", "Brems cross sections= NRC"))
    p, _ = parse(f)
    assert transport_options(p)["brems_xsec"] == "2"
    f.write_text(txt)
    p, _ = parse(f)
    assert transport_options(p)["brems_xsec"] == "0"

"""Electron impact ionisation ("Electron impact ionization = On": eii_init,
egsnrc.mortran:7434-7758; the EII branch of MOLLER, :2634-2664;
eii_sample, :7760-7830), with the tutorial media (Ta, Pb, water) of
egsnrc_amd/data/tutor_eii.egstab.

CPU: the blob's shell cross sections against data/eii_ik.data (fixture
tests/golden/eii_ik_ta_pb.json): the grid starts at the binding energy and
every bin ends on the tabulated value; the oracle's sampled energy transfers
have the mean that eii_init's closed-form integrals give for the same shell
(two independent expressions of one differential cross section); the option
and the blob must agree.  GPU: dose parity of the engine with the oracle
under 1 MeV electrons on lead and water."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from egsnrc_amd import _lib
from egsnrc_amd.engine import Simulation, Tables, make_source
from tests import oracle_lib
from tests.blob import read_blob
from tests.parity import check_parity

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BLOB = os.path.join(ROOT, "egsnrc_amd", "data", "tutor_eii.egstab")
BLOB_PLAIN = os.path.join(ROOT, "egsnrc_amd", "data", "tutor.egstab")
FIX = os.path.join(ROOT, "tests", "golden", "eii_ik_ta_pb.json")
RM = 0.5109989461
TA, PB, WATER = 0, 1, 2


def test_shell_cross_sections_follow_the_data_file():
    fx = json.load(open(FIX))
    b = read_blob(BLOB)
    bind = b["binding_energies"]
    nbin = fx["nbin"]
    assert int(b["eii"][0]) == 1
    for jj, (z, sh) in enumerate(zip(b["eii_z"], b["eii_sh"])):
        d = fx["shells"]["%d.%d" % (z, sh)]
        U = bind[int(z) * 30 + int(sh) - 1]
        a, bb = b["eii_a"][jj], b["eii_b"][jj]
        assert abs(a - nbin / np.log(fx["emax"] / U)) < 1e-9 * a
        assert abs(np.exp((1 - bb) / a) - U) < 1e-9 * U          # the grid starts at U
        k = np.arange(1, nbin + 1)
        loge = (k + 1 - bb) / a                                     # upper edge of bin k
        xa = b["eii_xa"][jj * nbin:(jj + 1) * nbin]
        xb = b["eii_xb"][jj * nbin:(jj + 1) * nbin]
        want = d["fmax"] * np.array(d["rel"])
        assert np.allclose(xa * loge + xb, want, rtol=1e-9, atol=1e-9 * d["fmax"]), (z, sh)
    # per medium: Ta's four shells first, then lead's; none in water
    assert list(b["m0.eii_first"]) == [1] and list(b["m1.eii_first"]) == [5]
    assert int(b["m2.eii_nsh"][0]) == 0


def av_transfer(e, U):
    """Mean energy transfer of an EII event on a shell of binding U at
    kinetic energy e: the ratio of the first to the zeroth moment of the
    soft and hard parts of the cross section, integrated in closed form
    (eii_init, egsnrc.mortran:7662-7676)."""
    tau = e / RM
    beta2 = tau * (tau + 2) / (tau + 1) ** 2
    p2 = 2 * RM * tau * (tau + 2)
    Wmax = (e + U) / 2
    uwm = U / Wmax
    ss_0 = 2 * (np.log(p2 / U) - uwm ** 3 * np.log(p2 / Wmax) - (beta2 + 0.833333) * (1 - uwm ** 3)) / 3 / U
    sh_0 = ((1 - uwm) * (1 + uwm / (2 - uwm)) + U * (Wmax - U) / (e + RM) ** 2
            - (2 * tau + 1) / (tau + 1) ** 2 * uwm / 2 * np.log((2 - uwm) / uwm)) / U
    ss_1 = np.log(p2 / U) - uwm ** 2 * np.log(p2 / Wmax) - (beta2 + 1) * (1 - uwm ** 2)
    sh_1 = (np.log(Wmax / U / (2 - uwm)) + 2 * (Wmax - U) / (2 * Wmax - U)
            + (Wmax ** 2 - U ** 2) / (e + RM) ** 2 / 2
            - (2 * tau + 1) / (tau + 1) ** 2 * np.log((2 * Wmax - U) / Wmax))
    return (ss_1 + sh_1) / (ss_0 + sh_0)


def test_sampled_transfer_has_the_integrated_mean():
    for med in (PB, TA):
        for ish in (1, 2, 4):
            for e in (0.2, 1.0, 10.0):
                out = oracle_lib.sample(BLOB, 15, med, e, 200000, param=ish)
                wx, U = out[:, 0], out[0, 3]
                assert np.all(wx >= U * (1 - 1e-12)) and np.all(wx <= (e + U) / 2 * (1 + 1e-9))
                m, se = wx.mean(), wx.std() / np.sqrt(len(wx))
                want = av_transfer(e, U)
                assert abs(m - want) < 4 * se + 1e-3 * want, (med, ish, e, m, want)
                # the secondary leaves with wx - U when above AE - m, else
                # the transfer less the binding is absorbed here
                sec = out[:, 1]
                assert np.allclose(sec[sec > 0], wx[sec > 0] - U, atol=1e-12)


def test_option_and_blob_must_agree():
    L = _lib.lib()
    xb = np.array([0.0, 1.0])
    g, keep = _lib.make_geometry(xb, xb.copy(), xb.copy(), np.zeros(1, np.int32), None)
    src = make_source(-1, 1.0)
    for path, eii in ((BLOB, b""), (BLOB, b"Off"), (BLOB_PLAIN, b"On")):
        tab = Tables(path)
        tp = _lib.default_transport(eii_xsec=eii)
        h = C.c_void_p()
        assert L.egsgpu_create(tab.h, C.byref(g), C.byref(src[0]), C.byref(tp), C.byref(h)) == -1
        buf = C.create_string_buffer(512)
        L.egsgpu_last_error(buf, 512)
        assert b"EII" in buf.value or b"eii" in buf.value


def eii_case(n=32, dx=0.0375, energy=1.0, half=0.6):
    """1 MeV electrons over the whole 1.2 x 1.2 cm^2 face of 1 mm of lead
    (0.1 mm layers, deeper than the electrons' range) over water, EII on
    (tutorial data: ECUT 0.7 MeV, PCUT 0.01 MeV)."""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.concatenate([np.linspace(0.0, 0.1, 11), np.linspace(0.2, 1.0, 9)])
    zc = 0.5 * (zb[1:] + zb[:-1])
    mz = np.where(zc < 0.1, PB, WATER)
    med = np.broadcast_to(mz[:, None, None], (len(zc), n, n)).reshape(-1).astype(np.int32)
    src = make_source(-1, energy, x0=-half, x1=half, y0=-half, y1=half, z=-1.0)
    tp = _lib.default_transport(kerma=0, eii_xsec=b"On")
    return xb, xb.copy(), zb, med, None, src, tp


HIST = 200000
_ORACLE = {}


def _oracle(seeds=None):
    if seeds not in _ORACLE:
        xb, yb, zb, med, rhor, src, tp = eii_case()
        g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
        o = oracle_lib.OracleSimulation(BLOB, g, src[0], tp, nstreams=256,
                                        nthreads=max(1, min(os.cpu_count() or 1, 16)))
        if seeds:
            o.seed(*seeds)
        o.run(HIST)
        _ORACLE[seeds] = o.tallies()
    return _ORACLE[seeds]


def test_oracle_eii_run_deposits_like_without():
    """EII redistributes the Moller events of the K and L shells; the
    energy lead absorbs from 1 MeV electrons stays within 2 % of the run
    without it."""
    xb, yb, zb, med, rhor, src, tp = eii_case()
    g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
    res = {}
    for name, path, t in (("eii", BLOB, tp), ("plain", BLOB_PLAIN, _lib.default_transport(kerma=0))):
        o = oracle_lib.OracleSimulation(path, g, src[0], t, nstreams=64, nthreads=8)
        o.run(20000)
        s, s2, n = o.tallies()
        res[name] = s.sum() / n
    print(res)
    assert abs(res["eii"] - res["plain"]) < 0.02 * res["plain"]


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["event", "history"])
def test_eii_dose_parity(engine, schedule):
    xb, yb, zb, med, rhor, src, tp = eii_case()
    sim = Simulation(Tables(BLOB), xb, yb, zb, med, rhor, src, tp, nstreams=8192,
                     schedule=schedule)
    sim.run(HIST)
    b = sim.tallies()
    assert b[2] == HIST
    c = sim.counters()
    assert c[5] == 0 and c[6] == 0
    check_parity(_oracle(), b, c, label="eii/" + schedule, null=_oracle((4242, 2718)))

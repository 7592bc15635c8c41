"""Pinning the oracle's interaction samplers against closed forms that are
derived independently of the reference's rejection functions (PENELOPE-style
or textbook forms), plus kinematic invariants.  A dropped term or wrong sign
in a sampler fails a chi^2 test here.  (CPU only.)"""
import os

import numpy as np
from scipy import integrate, stats

from tests import oracle_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BLOB = os.path.join(ROOT, "egsnrc_amd", "data", "521icru.egstab")
RM = 0.5109989461
WATER = 0


def chi2_test(samples, pdf, lo, hi, nbins=40):
    """chi^2 p-value of a histogram against an unnormalised pdf on [lo,hi]."""
    edges = np.linspace(lo, hi, nbins + 1)
    h, _ = np.histogram(samples, edges)
    p = np.array([integrate.quad(pdf, a, b, limit=200)[0] for a, b in zip(edges[:-1], edges[1:])])
    p = p / p.sum()
    exp = p * len(samples)
    m = exp > 5
    chi2 = np.sum((h[m] - exp[m]) ** 2 / exp[m])
    return stats.chi2.sf(chi2, m.sum() - 1)


def test_klein_nishina_energy_distribution():
    k = 1.0
    kap = k / RM
    out = oracle_lib.sample(BLOB, 0, WATER, k, 200000)   # param 0 -> ibcmp 0
    eps = out[:, 0]
    emin = 1 / (1 + 2 * kap)
    assert eps.min() >= emin - 1e-12 and eps.max() <= 1 + 1e-12

    def kn(e):   # d sigma / d eps (Klein-Nishina)
        c = 1 - (1 - e) / (kap * e)
        return (1 / e + e) * (1 - e * (1 - c * c) / (1 + e * e))
    assert chi2_test(eps, kn, emin, 1.0) > 1e-3
    # photon direction: cos theta = 1 - (1-eps)/(kap eps)
    c = 1 - (1 - eps) / (kap * eps)
    assert np.allclose(out[:, 1], c, atol=1e-9)


def test_bound_compton_energy_conservation():
    k = 0.1
    out = oracle_lib.sample(BLOB, 0, WATER, k, 50000, param=3.0)
    ok = out[:, 0] > -1
    kp = out[ok, 0] * k
    assert np.allclose(kp + out[ok, 2] + out[ok, 3], k, atol=1e-12)


def _penelope_moller(E):
    g = E / RM + 1
    a = ((g - 1) / g) ** 2

    def f(W):
        return (1 / W ** 2) * (1 + (W / (E - W)) ** 2 - (1 - a) * W / (E - W) + a * (W / E) ** 2)
    return f


def test_moller_secondary_energy_distribution():
    E = 5.0
    out = oracle_lib.sample(BLOB, 1, WATER, E, 200000)
    W = out[:, 0]
    te = 0.521 - RM
    assert W.min() >= te - 1e-9 and W.max() <= E / 2 + 1e-9
    assert np.allclose(out[:, 0] + out[:, 1], E, atol=1e-9)
    assert chi2_test(np.log(W), lambda lw: _penelope_moller(E)(np.exp(lw)) * np.exp(lw),
                     np.log(te), np.log(E / 2)) > 1e-3


def test_bhabha_energy_distribution():
    E = 3.0
    out = oracle_lib.sample(BLOB, 2, WATER, E, 200000)
    # the particle with energy fraction w of the kinetic energy
    t2, q2 = out[:, 0], out[:, 2]
    w = np.where(q2 < 0, t2, E - t2)     # kinetic energy of the electron
    te = 0.521 - RM
    g = E / RM + 1
    c = ((g - 1) / g) ** 2
    b1 = c * (2 * (g + 1) ** 2 - 1) / (g * g - 1)
    b2 = c * (3 * (g + 1) ** 2 + 1) / (g + 1) ** 2
    b3 = c * 2 * g * (g - 1) / (g + 1) ** 2
    b4 = c * (g - 1) ** 2 / (g + 1) ** 2

    def f(W):   # Bhabha (PENELOPE form), W = electron kinetic energy
        x = W / E
        return (1 / W ** 2) * (1 - b1 * x + b2 * x * x - b3 * x ** 3 + b4 * x ** 4)
    assert w.min() >= te - 1e-9
    assert chi2_test(np.log(w), lambda lw: f(np.exp(lw)) * np.exp(lw), np.log(te), np.log(E)) > 1e-3
    assert np.allclose(out[:, 0] + out[:, 1], E, atol=1e-9)


def test_annihilation_in_flight_heitler():
    E = 2.0
    out = oracle_lib.sample(BLOB, 3, WATER, E, 200000)
    g = E / RM + 1            # positron total energy / m
    etot = E + 2 * RM
    assert np.allclose(out[:, 0] + out[:, 1], etot, atol=1e-9)
    z = out[:, 0] / etot
    zmin = 1 / (g + 1 + np.sqrt(g * g - 1))

    def S(x):
        return -(g + 1) ** 2 + (g * g + 4 * g + 1) / x - 1 / (x * x)
    # either photon of the pair follows Heitler's symmetric distribution
    both = np.concatenate([z, 1 - z])
    assert chi2_test(both, lambda x: S(x) + S(1 - x), zmin, 1 - zmin) > 1e-3
    # two-body kinematics fixes the photon angle:
    # cos = (k - m) (E + 2m) / (k p) with p the positron momentum
    p = np.sqrt(E * (E + 2 * RM))
    c1 = (out[:, 0] - RM) * etot / (out[:, 0] * p)
    assert np.allclose(np.clip(c1, -1, 1), out[:, 2], atol=1e-9)


def test_pair_energy_conservation_and_symmetry():
    for k in (1.5, 6.0, 60.0):
        out = oracle_lib.sample(BLOB, 4, WATER, k, 40000)
        assert np.allclose(out[:, 0] + out[:, 1], k, atol=1e-9)
        assert np.all(out[:, 0] >= RM - 1e-12) and np.all(out[:, 1] >= RM - 1e-12)
        assert set(np.unique(out[:, 2] * out[:, 3])) == {-1.0}
        # the electron takes the larger share half of the time
        eminus = np.where(out[:, 2] < 0, out[:, 0], out[:, 1])
        frac = np.mean(eminus > k / 2)
        assert abs(frac - 0.5) < 0.02


def test_photo_energy_conservation():
    """Photo-electron plus the energy absorbed on the spot equals the photon
    energy, except for a vacancy in a shell without an XCOM edge (carbon's
    L3): PHOTO gives the electron the whole photon energy (its binding
    energy there is 0) and the relaxation then absorbs the EADL binding
    energy of the shell (8.98 eV), as egs_eadl_relax does."""
    for k in (0.02, 0.1, 1.0):
        out = oracle_lib.sample(BLOB, 5, 2, k, 20000)   # ICRP bone
        r = out[:, 0] - RM + out[:, 1] - k
        assert np.all((np.abs(r) < 1e-12) | (np.abs(r - 8.98e-6) < 1e-9)), r[np.abs(r) > 1e-12]
        assert np.mean(np.abs(r) > 1e-12) < 1e-3


def test_brems_kinematics():
    E = 10.0
    out = oracle_lib.sample(BLOB, 7, WATER, E, 50000)
    k = out[:, 0]
    assert k.min() >= 0.01 - 1e-12 and k.max() <= E + 1e-12
    assert np.allclose(out[:, 0] + out[:, 1], E + RM, atol=1e-9)
    # the angular distribution is forward peaked, <theta> ~ m/E
    th = np.arccos(np.clip(out[:, 2], -1, 1))
    assert np.median(th) < 5 * RM / E


def test_mscat_first_moment():
    """Goudsmit-Saunderson first moment: for screened Rutherford single
    scattering, pdf(t = 1 - cos) = 2 c(1+c)/(t+2c)^2 on [0,2] whose mean is
    q1 = 2c((1+c) log(1+1/c) - 1); after a Poisson number (mean lambda) of
    scatterings <cos> = exp(-lambda q1).  Checked across the no/single/few
    scattering branches and the q^(2+) surface."""
    for E, step in [(1.0, 0.002), (1.0, 0.02), (5.0, 0.2), (0.1, 0.005), (10.0, 1.0)]:
        out = oracle_lib.sample(BLOB, 6, WATER, E, 200000, param=step)
        lam, chia2 = out[0, 2], out[0, 3]
        q1 = 2 * chia2 * ((1 + chia2) * np.log(1 + 1 / chia2) - 1)
        expect = 1 - np.exp(-lam * q1)
        got = np.mean(1 - out[:, 0])
        err = np.std(1 - out[:, 0]) / np.sqrt(len(out))
        assert abs(got - expect) < max(5 * err, 0.02 * expect), (E, step, lam, got, expect)
        assert np.allclose(out[:, 0] ** 2 + out[:, 1] ** 2, 1, atol=1e-9)


def test_rayleigh_angles_forward_with_energy():
    m1 = np.mean(oracle_lib.sample(BLOB, 8, WATER, 0.02, 20000)[:, 0])
    m2 = np.mean(oracle_lib.sample(BLOB, 8, WATER, 0.5, 20000)[:, 0])
    assert -1 <= m1 < m2 <= 1
    assert m2 > 0.9

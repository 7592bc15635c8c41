"""Splitting and Russian roulette.

DOSXYZnrc photon splitting (n_split, dosxyznrc.mortran:855-1030): every
photon is replaced by n_split sub-photons of 1/n_split of its weight with
stratified interaction points; photons scattered by all but one randomly
chosen sub-photon are removed and the survivor's carry the full weight;
photons of bremsstrahlung and annihilation in flight or at rest play
Russian roulette (ausgab 7, 13, 14: :3531-3545, switched on at
:2557-2562).

Splitting is unbiased: the split run's expected dose is that of the
unsplit oracle, checked voxel by voxel against its statistics, with a
plainly wrong weight (no 1/n_split, or no n_split on a survivor) far
outside them.  Its variance per history is lower where the dose comes from
the photons' first interactions, which shows that the photons were split.

EGSnrc radiative splitting (nbr_split of BREMS, ANNIH and ANNIH_AT_REST,
egsnrc.mortran:70-526), with or without the Russian roulette of the charged
particles made by photon interactions (i_play_RR, egsnrc.macros:3481-3519;
PAIR's own roulette :4552-4577), as the egs_radiative_splitting ausgab
object sets them (egs_advanced_application.cpp:1224-1239): the same bar
against the unsplit oracle."""
import os
import subprocess

import numpy as np
import pytest

from egsnrc_amd import _lib
from egsnrc_amd.engine import Simulation, Tables, make_source
from egsnrc_amd import workloads
from tests import oracle_lib
from tests.parity import P2, check_parity, chi2_bound
from tests.test_host_app import APP, EX

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BLOB700 = os.path.join(ROOT, "egsnrc_amd", "data", "700icru.egstab")
BLOB521 = os.path.join(ROOT, "egsnrc_amd", "data", "521icru.egstab")


NULL_SEEDS = (4711, 20000)   # a second oracle run, independent of the first


def oracle(blob, xb, yb, zb, med, rho, src, tp, n, seeds=None):
    g, keep = _lib.make_geometry(xb, yb, zb, med, rho)
    o = oracle_lib.OracleSimulation(blob, g, src[0], tp, nstreams=256, nthreads=8)
    if seeds:
        o.seed(*seeds)
    o.run(n)
    return o.tallies()


def stats(t):
    s, s2, n = t
    m = s / n
    return m, np.maximum(s2 / n - m * m, 0) / (n - 1)


def check(a, b, label, frac=0.1, counters=None, null=None):
    check_parity(a, b, counters, label=label, frac=frac, null=null)


def per_history_variance_ratio(a, b, frac=0.1):
    """Median over the dosed voxels of b's variance per history over a's."""
    ma, va = stats(a)
    mb, vb = stats(b)
    keep = (ma > frac * ma.max()) & (va > 0) & (vb > 0)
    return float(np.median((vb[keep] * b[2]) / (va[keep] * a[2])))


@pytest.mark.gpu
def test_split_photon_beam_matches_the_unsplit_oracle():
    """1.25 MeV photons, 12 x 12 cm field, 16 x 16 x 20 water voxels of 1 cm
    with a bone slab at 5-7 cm; n_split 20, FP32 and ranmar."""
    xb = np.linspace(-8, 8, 17)
    zb = np.linspace(0, 20, 21)
    med = np.zeros(16 * 16 * 20, np.int32)
    kz = np.arange(med.size) // 256
    med[(kz == 5) | (kz == 6)] = 1
    src = make_source(0, 1.25, x0=-6, x1=6, y0=-6, y1=6, z=-1.0)
    tp = _lib.default_transport(ecut=0.7, pcut=0.01)
    a = oracle(BLOB700, xb, xb.copy(), zb, med, None, src, tp, 400000)
    a0 = oracle(BLOB700, xb, xb.copy(), zb, med, None, src, tp, 400000, NULL_SEEDS)
    tp.n_split = 20
    sim = Simulation(Tables(BLOB700), xb, xb.copy(), zb, med, None, src, tp, nstreams=16384)
    n = 100000
    sim.run(n)
    b = sim.tallies()
    c = sim.counters()
    assert b[2] == n and c[5] == 0
    check(a, b, "split photons", counters=c, null=a0)
    ratio = per_history_variance_ratio(a, b)
    print("variance per history, split / unsplit: %.3f" % ratio)
    assert ratio < 0.6


@pytest.mark.gpu
@pytest.mark.parametrize("charge,energy,fp64", [(-1, 10.0, 0), (1, 6.0, 1)])
def test_russian_roulette_of_radiative_photons(charge, energy, fp64):
    """Electrons (bremsstrahlung) and positrons (bremsstrahlung and
    annihilation in flight and at rest) on water slabs, n_split 10: the
    slabs beyond the particles' range see only photons that survived the
    roulette, and their dose must be the oracle's."""
    a, b, nz, bt = slab_run(charge, energy, fp64, 200000, n_split=10)
    check(a, b, "roulette q=%d" % charge, frac=0.05)
    tail_check(a, b, nz, "roulette q=%d" % charge, bt)


@pytest.mark.gpu
def test_split_with_howfarless_and_ranlux():
    """HOWFARLESS in homogeneous water with ranlux streams and n_split 10:
    photon splitting on the one-step photon flights of HOWFARLESS."""
    xb = np.linspace(-8, 8, 17)
    zb = np.linspace(0, 16, 17)
    med = np.zeros(16 ** 3, np.int32)
    src = make_source(0, 2.0, x0=-6, x1=6, y0=-6, y1=6, z=-1.0)
    tp = _lib.default_transport(ecut=0.7, pcut=0.01, rayleigh=0)
    a = oracle(BLOB700, xb, xb.copy(), zb, med, None, src, tp, 400000)
    a0 = oracle(BLOB700, xb, xb.copy(), zb, med, None, src, tp, 400000, NULL_SEEDS)
    tp.n_split = 10
    tp.howfarless = 1
    sim = Simulation(Tables(BLOB700), xb, xb.copy(), zb, med, None, src, tp, nstreams=16384)
    sim.set_ranlux(1, 7)
    n = 100000
    sim.run(n)
    b = sim.tallies()
    c = sim.counters()
    assert b[2] == n and c[5] == 0
    check(a, b, "split + howfarless + ranlux", counters=c, null=a0)
    assert per_history_variance_ratio(a, b) < 0.8


@pytest.mark.gpu
def test_split_rejects_other_schedules_and_the_gradient():
    xb = np.linspace(-2, 2, 5)
    med = np.zeros(64, np.int32)
    tp = _lib.default_transport(ecut=0.7, pcut=0.01, n_split=4)
    with pytest.raises(_lib.EngineError, match="plain history"):
        Simulation(Tables(BLOB700), xb, xb.copy(), xb + 2, med, None,
                   make_source(0, 1.0, z=-1.0), tp, schedule="event")
    tp.kerma = 1
    sim = Simulation(Tables(BLOB700), xb, xb.copy(), xb + 2, med, None,
                     make_source(0, 1.0, z=-1.0), tp)
    with pytest.raises(_lib.EngineError, match="unsplit"):
        sim.density_grad(1000, np.zeros(64))
    with pytest.raises(_lib.EngineError, match="0..256"):
        Simulation(Tables(BLOB700), xb, xb.copy(), xb + 2, med, None,
                   make_source(0, 1.0, z=-1.0), _lib.default_transport(n_split=300))


LAT = np.linspace(-6.0, 6.0, 25)   # 0.5 cm columns across a 12 x 12 cm phantom


def slab_run(charge, energy, fp64, n, nbatch=10, **split):
    """An 8 x 8 cm beam on a 12 x 12 cm water phantom of 0.5 cm columns cut
    along z into 0.5 cm slabs to 6 cm (scaled by energy / 10 MeV), then
    thicker ones to 22 cm; oracle run unsplit, engine run with the given
    splitting parameters, both in nbatch equal batches.  Returns the
    tallies of both, the number of slabs and the per-batch voxel sums
    (nbatch x nvox) of both."""
    zb = np.concatenate([np.arange(0, 6.5, 0.5), [7.0, 8.0, 10.0, 13.0, 17.0, 22.0]]) \
        * (energy / 10.0)
    nz = len(zb) - 1
    nl = len(LAT) - 1
    med = np.zeros(nz * nl * nl, np.int32)
    src = make_source(charge, energy, x0=-4, x1=4, y0=-4, y1=4, z=-1.0)
    tp = _lib.default_transport(ecut=0.521, pcut=0.01, fp64=fp64)
    g, keep = _lib.make_geometry(LAT, LAT.copy(), zb, med, None)
    o = oracle_lib.OracleSimulation(BLOB521, g, src[0], tp, nstreams=256, nthreads=8)
    for k, v in split.items():
        setattr(tp, k, v)
    sim = Simulation(Tables(BLOB521), LAT, LAT.copy(), zb, med, None, src, tp, nstreams=16384)
    batches = ([], [])
    prev = (0.0, 0.0)
    for k in range(nbatch):
        o.run(n // nbatch)
        sim.run(n // nbatch)
        cur = (o.tallies()[0], sim.tallies()[0])
        for i in range(2):
            batches[i].append(cur[i] - prev[i])
        prev = cur
    a, b = o.tallies(), sim.tallies()
    c = sim.counters()
    assert b[2] == n - n % nbatch and c[5] == 0 and c[6] == 0
    return a, b, nz, (np.array(batches[0]), np.array(batches[1]))


def tail_check(a, b, nz, label, batches, group=4):
    """The photon-only dose beyond the range of the charged particles (the
    last four slabs), on bins of group x group columns, with each bin's
    uncertainty from the spread of its batch means.  A split run's tail
    variance is a small fraction of the unsplit oracle's, so a per-voxel
    z would inherit the skew of the oracle's few-history tail voxels
    (single-sample t statistics of a skewed variable have E[t^2] > 1);
    batch means of bins are close to normal, and their Welch t statistics
    are mapped onto normal scores before the shared bar is applied.
    Returns the median ratio of the bins' batch variances, split over
    unsplit."""
    from scipy import stats as st
    nl = len(LAT) - 1
    nb = batches[0].shape[0]
    per = a[2] // nb

    def bins(x):
        x = x.reshape(nb, nz, nl, nl)[:, nz - 4:]
        x = x.reshape(nb, 4, nl // group, group, nl // group, group).sum(axis=(3, 5))
        return x.reshape(nb, -1) / per
    xa, xb = bins(batches[0]), bins(batches[1])
    ma, mb = xa.mean(axis=0), xb.mean(axis=0)
    va, vb = xa.var(axis=0, ddof=1) / nb, xb.var(axis=0, ddof=1) / nb
    keep = (ma > 0.05 * ma.max()) & (va + vb > 0)
    t = (ma[keep] - mb[keep]) / np.sqrt(va[keep] + vb[keep])
    nu = (va[keep] + vb[keep]) ** 2 / ((va[keep] ** 2 + vb[keep] ** 2) / (nb - 1))
    z = st.norm.ppf(np.clip(st.t.cdf(t, nu), 1e-300, 1 - 1e-16))
    chi2, out2, nv = float(np.mean(z * z)), float(np.mean(np.abs(z) > 2)), int(z.size)
    band = 3 * np.sqrt(P2 * (1 - P2) / nv)
    msg = ("%s tail bins: chi2/dof %.4f  >2sigma %.4f (allowed %.4f..%.4f)  bins %d  "
           "total ratio %.4f" % (label, chi2, out2, P2 - band, P2 + band, nv,
                                 mb[keep].sum() / ma[keep].sum()))
    print(msg)
    assert nv >= 100, msg
    assert chi2 <= chi2_bound(nv), msg
    assert abs(out2 - P2) <= band, msg
    return float(np.median(vb[keep] / va[keep]))


@pytest.mark.gpu
@pytest.mark.parametrize("play_rr", [1, 0])
def test_radiative_splitting_of_an_electron_beam(play_rr):
    """10 MeV electrons, nbr_split 20, with and without the roulette of the
    charged particles the split photons set in motion: the dose of every
    slab is the oracle's; without roulette the bremsstrahlung tail has a
    lower variance per history."""
    a, b, nz, bt = slab_run(-1, 10.0, 0, 200000, nbr_split=20, play_rr=play_rr)
    check(a, b, "radiative splitting rr=%d" % play_rr, frac=0.05)
    ratio = tail_check(a, b, nz, "rr=%d" % play_rr, bt)
    print("tail variance per history, split / unsplit: %.3f" % ratio)
    # without roulette every split photon's electrons are kept; with it
    # about as many electrons as without splitting survive, at full weight
    assert ratio < (0.5 if not play_rr else 2.0)


@pytest.mark.gpu
def test_radiative_splitting_of_annihilation():
    """6 MeV positrons in FP64 with nbr_split 8 and roulette: annihilation
    in flight and at rest make 8 pairs each."""
    a, b, nz, bt = slab_run(1, 6.0, 1, 200000, nbr_split=8, play_rr=1)
    check(a, b, "annihilation splitting", frac=0.05)
    tail_check(a, b, nz, "annihilation", bt)


@pytest.mark.gpu
def test_russian_roulette_of_photon_made_electrons():
    """2 MeV photons on the 16 x 16 x 20 cm phantom with a bone slab,
    nbr_split 4 with roulette: three in four Compton, photo- and pair
    electrons are removed, the others carry four times the weight."""
    xb = np.linspace(-8, 8, 17)
    zb = np.linspace(0, 20, 21)
    med = np.zeros(16 * 16 * 20, np.int32)
    kz = np.arange(med.size) // 256
    med[(kz == 5) | (kz == 6)] = 2
    src = make_source(0, 2.0, x0=-6, x1=6, y0=-6, y1=6, z=-1.0)
    tp = _lib.default_transport(ecut=0.521, pcut=0.01)
    a = oracle(BLOB521, xb, xb.copy(), zb, med, None, src, tp, 400000)
    a0 = oracle(BLOB521, xb, xb.copy(), zb, med, None, src, tp, 400000, NULL_SEEDS)
    tp.nbr_split = 4
    tp.play_rr = 1
    sim = Simulation(Tables(BLOB521), xb, xb.copy(), zb, med, None, src, tp, nstreams=16384)
    n = 800000
    sim.run(n)
    b = sim.tallies()
    assert b[2] == n and sim.counters()[5] == 0
    check(a, b, "electron roulette", counters=sim.counters(), null=a0)
    ratio = per_history_variance_ratio(a, b)
    print("variance per history, roulette / none: %.3f" % ratio)
    assert ratio > 1.5   # fewer, heavier electrons


@pytest.mark.gpu
def test_splitting_parameters_are_checked():
    xb = np.linspace(-2, 2, 5)
    med = np.zeros(64, np.int32)

    def make(**kw):
        Simulation(Tables(BLOB700), xb, xb.copy(), xb + 2, med, None,
                   make_source(0, 1.0, z=-1.0), _lib.default_transport(**kw))
    with pytest.raises(_lib.EngineError, match="choose one"):
        make(n_split=4, nbr_split=4)
    with pytest.raises(_lib.EngineError, match="needs nbr_split"):
        make(play_rr=1)
    with pytest.raises(_lib.EngineError, match="0..256"):
        make(nbr_split=1000)


def parse_transport(path):
    r = subprocess.run([APP, "-i", str(path), "--parse-only"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    line = [l for l in r.stdout.splitlines() if l.startswith("transport:")][0].split()[1:]
    return dict(zip(line[::2], line[1::2]))


@pytest.mark.parametrize("value,nbr,rr", [("20", "20", "1"), ("-10", "10", "0"),
                                          ("1", "0", "0")])
def test_app_reads_egs_radiative_splitting(tmp_path, value, nbr, rr):
    """'splitting = n' of an egs_radiative_splitting ausgab object
    (egs_radiative_splitting.h:110-121): n > 1 splits with roulette, n < -1
    without, 1 leaves the transport alone."""
    txt = open(EX).read().replace(
        ":stop ausgab object definition:",
        "    :start ausgab object:\n        name = vr\n"
        "        library = egs_radiative_splitting\n        splitting = %s\n"
        "    :stop ausgab object:\n:stop ausgab object definition:" % value)
    f = tmp_path / "vr.egsinp"
    f.write_text(txt)
    t = parse_transport(f)
    assert (t["nbr_split"], t["play_rr"]) == (nbr, rr)

"""Transport parity: the HIP engine (FP32 transport, FP64 tallies) against
the FP64 oracle on the same phantom, source and transport parameters.
Doses agree statistically voxel by voxel (independent arithmetic and
therefore independent histories): chi^2/dof near 1 and ~5% of voxels
outside 2 sigma, over voxels above 10% of Dmax."""
import os

import numpy as np
import pytest

from egsnrc_amd import _lib
from egsnrc_amd.engine import Simulation, Tables, make_source
from tests import oracle_lib, workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BLOB521 = os.path.join(ROOT, "egsnrc_amd", "data", "521icru.egstab")


def kerma_case(n=16, dx=1.6):
    xb, yb, zb, med, rhor = workloads.kerma_phantom(n, dx)
    src = make_source(0, 2.0, x0=-4, x1=4, y0=-4, y1=4, z=-1.0)
    tp = _lib.default_transport(ecut=20.0, kerma=1)
    return xb, yb, zb, med, rhor, src, tp


def test_oracle_kerma_energy_balance():
    """Every photon deposits between 0 and its energy; with a thick phantom
    the mean deposit is a sizeable fraction of 2 MeV."""
    xb, yb, zb, med, rhor, src, tp = kerma_case()
    g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
    o = oracle_lib.OracleSimulation(BLOB521, g, src[0], tp, nstreams=64, nthreads=4)
    o.run(20000)
    s, s2, n = o.tallies()
    assert n == 20000
    tot = s.sum() / n
    assert 0.5 < tot < 2.0
    # per-history deposit can never exceed the source energy
    assert np.all(s2 / n <= 4.0 * (s / n) + 1e-12)


def test_oracle_kerma_uniform_slab_attenuation():
    """Kerma in a homogeneous water slab behind the entrance follows the
    narrow-beam attenuation of primaries plus scatter; the first voxel
    layer receives the most dose per unit mass along a central pencil."""
    n = 1
    xb = np.array([-50.0, 50.0]); yb = xb.copy()
    zb = np.linspace(0, 40, 41)
    med = np.zeros(40, dtype=np.int32)
    g, keep = _lib.make_geometry(xb, yb, zb, med, None)
    src = make_source(0, 1.25, shape="point", x0=0, y0=0, z=-1.0)
    tp = _lib.default_transport(ecut=20.0, kerma=1, rayleigh=0)
    o = oracle_lib.OracleSimulation(BLOB521, g, src[0], tp, nstreams=64, nthreads=4)
    o.run(40000)
    s, s2, nh = o.tallies()
    d = s / nh
    # deposits fall off with depth after the build-up of scatter
    assert d[:5].mean() > d[-5:].mean()
    assert np.all(d > 0)


@pytest.mark.gpu
@pytest.mark.parametrize("fp64,schedule", [(0, "history"), (1, "history"), (0, "event")])
def test_kerma_dose_parity(engine, fp64, schedule):
    xb, yb, zb, med, rhor, src, tp = kerma_case()
    tp.fp64 = fp64
    g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
    o = oracle_lib.OracleSimulation(BLOB521, g, src[0], tp, nstreams=256, nthreads=8)
    o.run(400000)
    a = o.tallies()
    sim = Simulation(Tables(BLOB521), xb, yb, zb, med, rhor, src, tp, nstreams=16384,
                     schedule=schedule)
    sim.run(400000)
    b = sim.tallies()
    assert b[2] == 400000
    c = sim.counters()
    assert c[5] == 0, "stack overflow"
    chi2, out2, nv = workloads.compare_doses(a, b)
    print("chi2/dof %.3f  >2sigma %.3f  voxels %d" % (chi2, out2, nv))
    assert nv > 200
    assert chi2 < 1.3
    assert out2 < 0.09
    # total energy deposited agrees to within 4 combined sigma
    ta, tb = a[0].sum() / a[2], b[0].sum() / b[2]
    assert abs(ta - tb) < 0.01 * ta


@pytest.mark.gpu
@pytest.mark.parametrize("fp64,rng", [(0, "ranmar"), (1, "ranmar"), (0, "ranlux")])
def test_kerma_voting_kernel_runs_the_plain_histories(engine, fp64, rng):
    """Within one launch the kerma-mode history kernel with wave voting runs
    every lane's histories with the same numbers as the plain per-lane
    loop: identical event counters, and dose sums equal up to the order of
    the FP64 additions.  (Across launches the two may differ: numbers left
    in a lane's ring when a launch ends are dropped, and how many depends
    on when the wave refilled.)"""
    xb, yb, zb, med, rhor, src, tp = kerma_case()
    tp.fp64 = fp64
    out = []
    for schedule in ("history", "history-plain"):
        sim = Simulation(Tables(BLOB521), xb, yb, zb, med, rhor, src, tp, nstreams=8192,
                         schedule=schedule)
        if rng == "ranlux":
            sim.set_ranlux(4, 1)
        sim.run(200001)
        out.append((sim.tallies(), sim.counters()))
        sim.close()
    (a, ca), (b, cb) = out
    assert a[2] == b[2] == 200001
    assert list(ca) == list(cb)
    assert ca[2] > 0 and ca[1] > 0
    np.testing.assert_allclose(a[0], b[0], rtol=1e-12, atol=0)
    np.testing.assert_allclose(a[1], b[1], rtol=1e-12, atol=0)


def electron_case(n=12, dx=1.0, energy=6.0):
    """Electron beam on a water phantom with a bone slab (coupled transport,
    ECUT 0.521 MeV total, PCUT 0.01 MeV)."""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    zc = 0.5 * (zb[1:] + zb[:-1])
    mz = np.where((zc > 1.0) & (zc < 2.0), 2, 0)
    med = np.broadcast_to(mz[:, None, None], (n, n, n)).reshape(-1).astype(np.int32)
    src = make_source(-1, energy, x0=-2, x1=2, y0=-2, y1=2, z=-1.0)
    tp = _lib.default_transport(kerma=0)
    return xb, xb.copy(), zb, med, None, src, tp


def test_oracle_electron_depth_dose_20mev():
    """20 MeV broad electron beam in water: R50 ~ 8.3 cm (E0 ~ 2.33 R50)
    and energy deposited below E0 by the escaping bremsstrahlung."""
    xb = np.array([-20.0, 20.0]); yb = xb.copy(); zb = np.linspace(0, 15, 31)
    med = np.zeros(30, dtype=np.int32)
    g, keep = _lib.make_geometry(xb, yb, zb, med, None)
    src = make_source(-1, 20.0, x0=-5, x1=5, y0=-5, y1=5, z=-1.0)
    tp = _lib.default_transport(kerma=0)
    o = oracle_lib.OracleSimulation(BLOB521, g, src[0], tp, nstreams=64, nthreads=8)
    o.run(3000)
    s, s2, n = o.tallies()
    d = s / n / 0.5
    assert 17.5 < d.sum() * 0.5 < 19.5
    i = int(np.argmax(d))
    zc = zb[:-1] + 0.25
    r50 = np.interp(0.5 * d.max(), d[i:][::-1], zc[i:][::-1])
    assert 7.9 < r50 < 8.8, r50


def check_parity(a, b, c, min_voxels, label):
    """Statistical dose parity of two runs (per-voxel z-scores over voxels
    above 10% of the maximum) and no overflow of any engine structure."""
    chi2, out2, nv = workloads.compare_doses(a, b)
    print("%s: chi2/dof %.3f  >2sigma %.3f  voxels %d  overflows %d/%d"
          % (label, chi2, out2, nv, c[5], c[6]))
    assert c[5] == 0, "stack or bank overflow"
    assert c[6] == 0, "history tally table overflow"
    assert nv > min_voxels
    # 3-sigma bounds of chi2/dof and of the 2-sigma exceedance fraction for
    # nv independent normal z-scores (never tighter than 1.3 and 9 %)
    assert chi2 < max(1.3, 1 + 3 * np.sqrt(2.0 / nv))
    assert out2 < max(0.09, 0.0455 + 3 * np.sqrt(0.0455 * 0.9545 / nv))
    # energy deposited per history: ~4 sigma for 1e5 photon histories
    ta, tb = a[0].sum() / a[2], b[0].sum() / b[2]
    assert abs(ta - tb) < 0.02 * ta


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["history", "event"])
def test_electron_dose_parity(engine, schedule):
    xb, yb, zb, med, rhor, src, tp = electron_case()
    g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
    o = oracle_lib.OracleSimulation(BLOB521, g, src[0], tp, nstreams=256, nthreads=8)
    o.run(40000)
    a = o.tallies()
    sim = Simulation(Tables(BLOB521), xb, yb, zb, med, rhor, src, tp, nstreams=8192,
                     schedule=schedule)
    sim.run(20000)
    sim.run(20000)
    b = sim.tallies()
    assert b[2] == 40000
    check_parity(a, b, sim.counters(), 30, "electrons/" + schedule)


def photon_coupled_case(n=16, dx=1.0):
    """Photon spectrum (0.1-6 MeV, falling) on water/lung/bone/water slabs
    with coupled electron transport: every particle type and interaction
    of the slab workload at a size the oracle runs in seconds."""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    zc = 0.5 * (zb[1:] + zb[:-1])
    mz = np.select([zc < 4, zc < 8, zc < 10], [0, 1, 2], 0)
    med = np.broadcast_to(mz[:, None, None], (n, n, n)).reshape(-1).astype(np.int32)
    e, w = workloads.falling_spectrum()
    src = make_source(0, 0.0, spectrum=(e, w), x0=-3, x1=3, y0=-3, y1=3, z=-1.0)
    tp = _lib.default_transport(ecut=0.521, pcut=0.01, kerma=0)
    return xb, xb.copy(), zb, med, None, src, tp


_ORACLE_RUNS = {}


def _oracle_photon_coupled(n):
    """The oracle's tallies for photon_coupled_case (one run per n)."""
    if n not in _ORACLE_RUNS:
        xb, yb, zb, med, rhor, src, tp = photon_coupled_case()
        g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
        o = oracle_lib.OracleSimulation(BLOB521, g, src[0], tp, nstreams=256, nthreads=8)
        o.run(n)
        _ORACLE_RUNS[n] = o.tallies()
    return _ORACLE_RUNS[n]


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["history", "event", "event-host"])
def test_photon_coupled_dose_parity(engine, schedule):
    """event: iterations replayed from a HIP graph (the default);
    event-host: the same kernels launched one by one from the host."""
    xb, yb, zb, med, rhor, src, tp = photon_coupled_case()
    a = _oracle_photon_coupled(400000)
    sim = Simulation(Tables(BLOB521), xb, yb, zb, med, rhor, src, tp, nstreams=8192,
                     schedule=schedule.split("-")[0])
    if schedule == "event-host":
        _lib.check(_lib.lib().egsgpu_set_event_graph(sim.h, 0))
    sim.run(400000)
    b = sim.tallies()
    assert b[2] == 400000
    check_parity(a, b, sim.counters(), 100, "photons coupled/" + schedule)


@pytest.mark.gpu
@pytest.mark.parametrize("schedule,lux", [("history", 1), ("event", 1), ("event", 4)])
def test_ranlux_driven_transport_parity(engine, schedule, lux):
    """Coupled photon transport driven by ranlux streams (DOSXYZnrc's
    generator) agrees with the oracle's ranmar run."""
    xb, yb, zb, med, rhor, src, tp = photon_coupled_case()
    a = _oracle_photon_coupled(400000)
    sim = Simulation(Tables(BLOB521), xb, yb, zb, med, rhor, src, tp, nstreams=8192,
                     schedule=schedule)
    sim.set_ranlux(lux, 33)
    sim.run(400000)
    b = sim.tallies()
    assert b[2] == 400000
    check_parity(a, b, sim.counters(), 100, "ranlux %d %s" % (lux, schedule))


@pytest.mark.gpu
@pytest.mark.parametrize("rng", ["ranmar", "ranlux"])
def test_checkpoint_restart_continues_the_run(engine, rng):
    """A run restored from a checkpoint continues with the same random
    numbers: checkpoint after N1 histories, then N2 more, equals N1 + N2
    in one go (history schedule; FP64 atomics make the sums exact only up
    to their order)."""
    xb, yb, zb, med, rhor, src, tp = kerma_case()
    tab = Tables(BLOB521)
    a = Simulation(tab, xb, yb, zb, med, rhor, src, tp, nstreams=4096, schedule="history")
    if rng == "ranlux":
        a.set_ranlux(2, 77)
    a.run(20000)
    ck = a.checkpoint()
    a.run(30000)
    sa = a.tallies()
    b = Simulation(tab, xb, yb, zb, med, rhor, src, tp, nstreams=4096, schedule="history")
    b.restore(ck)
    b.run(30000)
    sb = b.tallies()
    assert sa[2] == sb[2] == 50000
    np.testing.assert_allclose(sb[0], sa[0], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(sb[1], sa[1], rtol=1e-12, atol=1e-12)
    # a checkpoint does not fit a simulation with other streams
    c = Simulation(tab, xb, yb, zb, med, rhor, src, tp, nstreams=2048, schedule="history")
    c.run(10)
    with pytest.raises(_lib.EngineError):
        c.restore(ck)


@pytest.mark.gpu
def test_fp32_vs_fp64_precision_report(engine):
    """FP32 transport against FP64 transport with the same random numbers:
    per-voxel differences stay well inside the statistical uncertainty, and
    the total deposit agrees within 3 sigma of the total's uncertainty
    (from the voxel uncertainties as if independent, which the positive
    correlation of a history's deposits only makes smaller).  Once a
    rounding sends a history another way, the two runs continue on
    different paths, so the total's difference is partly statistical and
    moves with any change to how the kernels are compiled."""
    from egsnrc_amd.precision_report import precision_report
    xb, yb, zb, med, rhor, src, tp = photon_coupled_case()
    rep, arrays = precision_report(Tables(BLOB521), xb, yb, zb, med, rhor, src, tp, 100000,
                                   nstreams=4096)
    print(rep)
    assert rep["voxels"] > 100
    assert rep["chi2_per_dof"] < 1.0
    assert rep["frac_beyond_2sigma"] < 0.05
    r32, s32, r64, s64 = arrays
    sig_total = np.sqrt(np.sum(s32 ** 2) + np.sum(s64 ** 2)) / r64.sum()
    print("total: rel diff %.5f, sigma %.5f" % (rep["total_rel_diff"], sig_total))
    assert abs(rep["total_rel_diff"]) < 3 * sig_total


def water_point_case(n=10, dx=1.5, energy=2.0):
    """A second geometry for the batch: water cube, 2 MeV point beam."""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    med = np.zeros(n ** 3, dtype=np.int32)
    src = make_source(0, energy, shape="point", x0=0.0, y0=0.0, z=-1.0)
    tp = _lib.default_transport(ecut=0.521, pcut=0.01, kerma=0)
    return xb, xb.copy(), zb, med, None, src, tp


@pytest.mark.gpu
def test_batched_jobs_match_oracle(engine):
    """Three simulations of different geometry and source run as one batch:
    each job's dose agrees with the oracle run of that job alone."""
    from egsnrc_amd.engine import run_batched
    cases = [electron_case(), photon_coupled_case(), water_point_case()]
    nh = [30000, 200000, 100000]
    tab = Tables(BLOB521)
    sims = []
    for c in cases:
        xb, yb, zb, med, rhor, src, tp = c
        tp.ecut, tp.pcut = 0.521, 0.01
        sims.append(Simulation(tab, xb, yb, zb, med, rhor, src, tp, nstreams=8192,
                               schedule="event"))
    run_batched(sims, nh)
    for c, sim, n, label in zip(cases, sims, nh, ["electron", "photon slabs", "point"]):
        xb, yb, zb, med, rhor, src, tp = c
        g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
        o = oracle_lib.OracleSimulation(BLOB521, g, src[0], tp, nstreams=256, nthreads=8)
        o.run(n)
        b = sim.tallies()
        assert b[2] == n
        check_parity(o.tallies(), b, sims[0].counters(), 30, "batch/" + label)
    # jobs must share transport parameters
    xb, yb, zb, med, rhor, src, tp = water_point_case()
    tp.ecut = 0.7
    other = Simulation(tab, xb, yb, zb, med, rhor, src, tp, nstreams=8192, schedule="event")
    with pytest.raises(_lib.EngineError):
        run_batched([sims[0], other], [10, 10])


@pytest.mark.gpu
def test_autotune_leaves_tallies_alone(engine):
    xb, yb, zb, med, rhor, src, tp = photon_coupled_case()
    sim = Simulation(Tables(BLOB521), xb, yb, zb, med, rhor, src, tp, nstreams=8192,
                     schedule="event")
    sim.run(20000)
    before = sim.tallies()
    c0 = sim.counters()
    choice = sim.autotune(20000)
    after = sim.tallies()
    assert after[2] == before[2] == 20000
    np.testing.assert_array_equal(after[0], before[0])
    np.testing.assert_array_equal(after[1], before[1])
    np.testing.assert_array_equal(sim.counters(), c0)
    wp, we, defer, batch, ks = choice
    assert wp in (2, 3) and we == wp and defer in (16, 32, 48) and batch in (8, 16, 24)
    assert ks in (4, 8, 16)


BLOB700 = os.path.join(os.path.dirname(BLOB521), "700icru.egstab")


@pytest.mark.gpu
def test_c0_pencil_beam_parity(engine):
    """C0 exactly: 1.25 MeV photon pencil beam on the centre of 30^3 x 1 cm
    water, ECUT 0.7 MeV (700icru data), PCUT 0.01 MeV, Rayleigh off."""
    xb, yb, zb, med, rhor = workloads.water_cube(30, 1.0)
    src = make_source(0, 1.25, shape="point", x0=0.0, y0=0.0, z=-1.0)
    tp = _lib.default_transport(ecut=0.7, pcut=0.01, rayleigh=0)
    g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
    o = oracle_lib.OracleSimulation(BLOB700, g, src[0], tp, nstreams=256, nthreads=8)
    o.run(500000)
    sim = Simulation(Tables(BLOB700), xb, yb, zb, med, rhor, src, tp)
    sim.run(500000)
    check_parity(o.tallies(), sim.tallies(), sim.counters(), 30, "c0")


@pytest.mark.gpu
def test_c2_electron_beam_reduced_parity(engine):
    """C2 at reduced size: 20 MeV electrons, 10x10 cm^2, on 64^3 x 0.4 cm
    water with a 3.2 cm aluminium cube at 3.2 cm depth."""
    xb, yb, zb, med, rhor = workloads.al_cube_phantom(64, 0.4, side=3.2, depth=3.2)
    src = make_source(-1, 20.0, x0=-5, x1=5, y0=-5, y1=5, z=-1.0)
    tp = _lib.default_transport(ecut=0.521, pcut=0.01)
    g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
    o = oracle_lib.OracleSimulation(BLOB521, g, src[0], tp, nstreams=256, nthreads=8)
    o.run(20000)
    sim = Simulation(Tables(BLOB521), xb, yb, zb, med, rhor, src, tp, nstreams=16384)
    sim.run(20000)
    check_parity(o.tallies(), sim.tallies(), sim.counters(), 100, "c2 reduced")

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
from egsnrc_amd import workloads
from tests import oracle_lib
from tests.parity import check_parity

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BLOB521 = os.path.join(ROOT, "egsnrc_amd", "data", "521icru.egstab")


def oracle_threads():
    return max(1, min(os.cpu_count() or 1, 16))


def kerma_case(n=32, dx=0.8):
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


KERMA_HIST = 4000000


@pytest.mark.gpu
@pytest.mark.parametrize("fp64,schedule", [(0, "history"), (1, "history"), (0, "event")])
def test_kerma_dose_parity(engine, fp64, schedule):
    xb, yb, zb, med, rhor, src, tp = kerma_case()
    tp.fp64 = fp64
    g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
    o = oracle_lib.OracleSimulation(BLOB521, g, src[0], tp, nstreams=256, nthreads=8)
    o.run(KERMA_HIST)
    a = o.tallies()
    sim = Simulation(Tables(BLOB521), xb, yb, zb, med, rhor, src, tp, nstreams=16384,
                     schedule=schedule)
    sim.run(KERMA_HIST)
    b = sim.tallies()
    assert b[2] == KERMA_HIST
    check_parity(a, b, sim.counters(), label="kerma/%s fp64=%d" % (schedule, fp64),
                 total_rtol=0.005)


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


def electron_case(n=30, dx=0.4, energy=6.0, half=4.0):
    """Electron beam (2*half x 2*half cm^2) on a water phantom with a bone
    slab at 1-2 cm depth (coupled transport, ECUT 0.521 MeV total, PCUT
    0.01 MeV)."""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    zc = 0.5 * (zb[1:] + zb[:-1])
    mz = np.where((zc > 1.0) & (zc < 2.0), 2, 0)
    med = np.broadcast_to(mz[:, None, None], (n, n, n)).reshape(-1).astype(np.int32)
    src = make_source(-1, energy, x0=-half, x1=half, y0=-half, y1=half, z=-1.0)
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


ELECTRON_HIST = 300000


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["history", "event"])
def test_electron_dose_parity(engine, schedule):
    xb, yb, zb, med, rhor, src, tp = electron_case()
    g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
    a = _oracle_cached("electron", electron_case, ELECTRON_HIST)
    sim = Simulation(Tables(BLOB521), xb, yb, zb, med, rhor, src, tp, nstreams=8192,
                     schedule=schedule)
    sim.run(ELECTRON_HIST // 2)
    sim.run(ELECTRON_HIST - ELECTRON_HIST // 2)
    b = sim.tallies()
    assert b[2] == ELECTRON_HIST
    check_parity(a, b, sim.counters(), label="electrons/" + schedule,
                 null=_oracle_cached("electron", electron_case, ELECTRON_HIST, NULL_SEEDS))


def lung_electron_case(n=32, dx=0.1, energy=2.0):
    """2 MeV electrons (1.6 x 1.6 cm^2) on a cube of lung (0.26 g/cm^3) in
    1 mm voxels.  At this density the elastic mean free path is long, so a
    large share of the steps are single-scattering steps near a voxel face,
    and many of them end before the face: the step branch whose energy loss
    is computed apart from the multiple-scattering and boundary steps
    (egsnrc.mortran:1444-1452).  Leaving that loss out moves 1.5 % of the
    energy deposited to greater depths."""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    med = np.ones(n ** 3, dtype=np.int32)      # lung of the 521icru blob
    src = make_source(-1, energy, x0=-0.8, x1=0.8, y0=-0.8, y1=0.8, z=-1.0)
    tp = _lib.default_transport(ecut=0.521, pcut=0.01, kerma=0)
    return xb, xb.copy(), zb, med, None, src, tp


LUNG_HIST = 400000


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["history", "event"])
def test_electron_beam_in_lung_parity(engine, schedule):
    """Voxel parity and the total deposit per history within 0.3 % (the
    oracle's own two runs differ by 2e-4, a missing single-scattering
    energy loss by 1.5e-2)."""
    xb, yb, zb, med, rhor, src, tp = lung_electron_case()
    a = _oracle_cached("lung electrons", lung_electron_case, LUNG_HIST)
    sim = Simulation(Tables(BLOB521), xb, yb, zb, med, rhor, src, tp, nstreams=8192,
                     schedule=schedule)
    sim.run(LUNG_HIST)
    b = sim.tallies()
    assert b[2] == LUNG_HIST
    check_parity(a, b, sim.counters(), label="lung electrons/" + schedule, total_rtol=3e-3,
                 null=_oracle_cached("lung electrons", lung_electron_case, LUNG_HIST,
                                     NULL_SEEDS))


def photon_coupled_case(n=32, dx=0.5):
    """Photon spectrum (0.1-6 MeV, falling) on water/lung/bone/water slabs
    with coupled electron transport: every particle type and interaction
    of the slab workload at a size the oracle runs in seconds."""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    zc = 0.5 * (zb[1:] + zb[:-1])
    mz = np.select([zc < 4, zc < 8, zc < 10], [0, 1, 2], 0)
    med = np.broadcast_to(mz[:, None, None], (n, n, n)).reshape(-1).astype(np.int32)
    e, w = workloads.falling_spectrum()
    src = make_source(0, 0.0, spectrum=(e, w), x0=-4, x1=4, y0=-4, y1=4, z=-1.0)
    tp = _lib.default_transport(ecut=0.521, pcut=0.01, kerma=0)
    return xb, xb.copy(), zb, med, None, src, tp


PHOTON_HIST = 2000000
_ORACLE_RUNS = {}


NULL_SEEDS = (4711, 20000)   # a second oracle run, independent of the first


def _oracle_cached(key, case, n, seeds=None):
    """The oracle's tallies for case() at n histories (one run per key, n
    and seeds, shared by the tests that compare against it)."""
    if (key, n, seeds) not in _ORACLE_RUNS:
        xb, yb, zb, med, rhor, src, tp = case()
        tp.ecut, tp.pcut = max(tp.ecut, 0.521), max(tp.pcut, 0.01)
        g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
        o = oracle_lib.OracleSimulation(BLOB521, g, src[0], tp, nstreams=256,
                                        nthreads=oracle_threads())
        if seeds:
            o.seed(*seeds)
        o.run(n)
        _ORACLE_RUNS[(key, n, seeds)] = o.tallies()
    return _ORACLE_RUNS[(key, n, seeds)]


def _oracle_photon_coupled(n):
    return _oracle_cached("photon slabs", photon_coupled_case, n)


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["history", "event", "event-host"])
def test_photon_coupled_dose_parity(engine, schedule):
    """event: iterations replayed from a HIP graph (the default);
    event-host: the same kernels launched one by one from the host."""
    xb, yb, zb, med, rhor, src, tp = photon_coupled_case()
    a = _oracle_photon_coupled(PHOTON_HIST)
    sim = Simulation(Tables(BLOB521), xb, yb, zb, med, rhor, src, tp, nstreams=8192,
                     schedule=schedule.split("-")[0])
    if schedule == "event-host":
        _lib.check(_lib.lib().egsgpu_set_event_graph(sim.h, 0))
    sim.run(PHOTON_HIST)
    b = sim.tallies()
    assert b[2] == PHOTON_HIST
    check_parity(a, b, sim.counters(), label="photons coupled/" + schedule,
                 null=_oracle_cached("photon slabs", photon_coupled_case, PHOTON_HIST, NULL_SEEDS))


@pytest.mark.gpu
@pytest.mark.parametrize("schedule,lux", [("history", 1), ("event", 1), ("event", 4)])
def test_ranlux_driven_transport_parity(engine, schedule, lux):
    """Coupled photon transport driven by ranlux streams (DOSXYZnrc's
    generator) agrees with the oracle's ranmar run."""
    xb, yb, zb, med, rhor, src, tp = photon_coupled_case()
    a = _oracle_photon_coupled(PHOTON_HIST)
    sim = Simulation(Tables(BLOB521), xb, yb, zb, med, rhor, src, tp, nstreams=8192,
                     schedule=schedule)
    sim.set_ranlux(lux, 33)
    sim.run(PHOTON_HIST)
    b = sim.tallies()
    assert b[2] == PHOTON_HIST
    check_parity(a, b, sim.counters(), label="ranlux %d %s" % (lux, schedule),
                 null=_oracle_cached("photon slabs", photon_coupled_case, PHOTON_HIST, NULL_SEEDS))


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
    the total deposit agrees within the uncertainty of its difference,
    estimated from ten paired batches (batch k of both runs holds the same
    histories; once a rounding sends a history another way the two runs
    continue on different paths, so the difference is partly statistical).
    The bound is the 0.1 % two-sided quantile of Student's t with 9 degrees
    of freedom (4.78)."""
    from egsnrc_amd.precision_report import precision_report
    xb, yb, zb, med, rhor, src, tp = photon_coupled_case()
    rep, arrays = precision_report(Tables(BLOB521), xb, yb, zb, med, rhor, src, tp, 400000,
                                   nstreams=4096)
    print(rep)
    assert rep["voxels"] > 1800
    assert rep["chi2_per_dof"] < 1.0
    assert rep["frac_beyond_2sigma"] < 0.0455
    print("total: rel diff %.5f, sigma %.5f" % (rep["total_rel_diff"],
                                               rep["total_rel_diff_sigma"]))
    assert abs(rep["total_rel_diff"]) < 4.78 * rep["total_rel_diff_sigma"]


def water_beam_case(n=24, dx=0.6, energy=2.0):
    """A third geometry for the batch: water cube, 2 MeV 6x6 cm^2 beam."""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    med = np.zeros(n ** 3, dtype=np.int32)
    src = make_source(0, energy, x0=-3.0, x1=3.0, y0=-3.0, y1=3.0, z=-1.0)
    tp = _lib.default_transport(ecut=0.521, pcut=0.01, kerma=0)
    return xb, xb.copy(), zb, med, None, src, tp


@pytest.mark.gpu
def test_batched_jobs_match_oracle(engine):
    """Three simulations of different geometry and source run as one batch:
    each job's dose agrees with the oracle run of that job alone."""
    from egsnrc_amd.engine import run_batched
    cases = [electron_case(), photon_coupled_case(), water_beam_case()]
    nh = [ELECTRON_HIST, PHOTON_HIST, 1000000]
    tab = Tables(BLOB521)
    sims = []
    for c in cases:
        xb, yb, zb, med, rhor, src, tp = c
        tp.ecut, tp.pcut = 0.521, 0.01
        sims.append(Simulation(tab, xb, yb, zb, med, rhor, src, tp, nstreams=8192,
                               schedule="event"))
    run_batched(sims, nh)
    for case, sim, n, label in zip((electron_case, photon_coupled_case, water_beam_case), sims,
                                   nh, ["electron", "photon slabs", "water beam"]):
        b = sim.tallies()
        assert b[2] == n
        check_parity(_oracle_cached(label, case, n), b, sims[0].counters(),
                     label="batch/" + label, null=_oracle_cached(label, case, n, NULL_SEEDS))
    # jobs must share transport parameters
    xb, yb, zb, med, rhor, src, tp = water_beam_case()
    tp.ecut = 0.7
    other = Simulation(tab, xb, yb, zb, med, rhor, src, tp, nstreams=8192, schedule="event")
    with pytest.raises(_lib.EngineError):
        run_batched([sims[0], other], [10, 10])


@pytest.mark.gpu
def test_autotune_leaves_tallies_alone(engine):
    """The probes leave tallies, history count and counters as they were;
    the choices are among the offered ones; the tuned schedule passes the
    coupled-photon parity."""
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
    wp, we, defer, batch, ks, spl, wg = choice
    assert wp in (2, 3) and we == wp and defer in (16, 32, 48) and batch in (8, 16, 24)
    assert ks in (4, 8, 16) and spl in (2, 4, 8) and wg in (128, 256)
    print("autotuned:", choice)
    # the tuned schedule (queue capacity and workgroup size included) still
    # transports correctly
    _lib.check(_lib.lib().egsgpu_reset_tallies(sim.h))
    sim.run(PHOTON_HIST)
    check_parity(_oracle_photon_coupled(PHOTON_HIST), sim.tallies(), sim.counters(),
                 label="photons coupled/autotuned",
                 null=_oracle_cached("photon slabs", photon_coupled_case, PHOTON_HIST, NULL_SEEDS))


BLOB700 = os.path.join(os.path.dirname(BLOB521), "700icru.egstab")


@pytest.mark.gpu
def test_c0_pencil_beam_parity(engine):
    """C0 exactly: 1.25 MeV photon pencil beam on the centre of 30^3 x 1 cm
    water, ECUT 0.7 MeV (700icru data), PCUT 0.01 MeV, Rayleigh off.  A
    pencil beam puts 10 % of Dmax in about 120 voxels, too few for the
    chi^2 bar to mean anything; the comparison is over the voxels above
    0.2 % of Dmax (about 2700 at 6e6 histories)."""
    xb, yb, zb, med, rhor = workloads.water_cube(30, 1.0)
    src = make_source(0, 1.25, shape="point", x0=0.0, y0=0.0, z=-1.0)
    tp = _lib.default_transport(ecut=0.7, pcut=0.01, rayleigh=0)
    g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
    n = 6000000
    ref = []
    for seeds in (None, NULL_SEEDS):
        o = oracle_lib.OracleSimulation(BLOB700, g, src[0], tp, nstreams=256,
                                        nthreads=oracle_threads())
        if seeds:
            o.seed(*seeds)
        o.run(n)
        ref.append(o.tallies())
        del o
    sim = Simulation(Tables(BLOB700), xb, yb, zb, med, rhor, src, tp)
    sim.run(n)
    check_parity(ref[0], sim.tallies(), sim.counters(), label="c0", frac=0.002,
                 total_rtol=0.003, null=ref[1])


@pytest.mark.gpu
def test_c2_electron_beam_reduced_parity(engine):
    """C2 at reduced size: 20 MeV electrons, 10x10 cm^2, on 64^3 x 0.4 cm
    water with a 3.2 cm aluminium cube at 3.2 cm depth."""
    xb, yb, zb, med, rhor = workloads.al_cube_phantom(64, 0.4, side=3.2, depth=3.2)
    src = make_source(-1, 20.0, x0=-5, x1=5, y0=-5, y1=5, z=-1.0)
    tp = _lib.default_transport(ecut=0.521, pcut=0.01)
    g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
    o = oracle_lib.OracleSimulation(BLOB521, g, src[0], tp, nstreams=256, nthreads=8)
    o.run(100000)
    # a second oracle run calibrates the 2-sigma fraction of these
    # heavy-tailed voxel errors (the Al cube's low-dose periphery)
    on = oracle_lib.OracleSimulation(BLOB521, g, src[0], tp, nstreams=256, nthreads=8)
    on.seed(*NULL_SEEDS)
    on.run(100000)
    sim = Simulation(Tables(BLOB521), xb, yb, zb, med, rhor, src, tp, nstreams=16384)
    sim.run(100000)
    check_parity(o.tallies(), sim.tallies(), sim.counters(), label="c2 reduced",
                 null=on.tallies())


@pytest.mark.gpu
def test_c2_full_size_parity(engine):
    """C2 at full size: 20 MeV electrons, 10 x 10 cm^2, on 256^3 x 0.2 cm
    water with a 3 cm aluminium cube at 3 cm depth (spin effects off: the
    reference tree has no spin data), 2e5 histories on the default event
    schedule against the oracle, with a second oracle run for the 2-sigma
    fraction (a voxel of this phantom gets its dose from about 80
    histories)."""
    xb, yb, zb, med, rhor = workloads.al_cube_phantom(256, 0.2)
    src = make_source(-1, 20.0, x0=-5, x1=5, y0=-5, y1=5, z=-1.0)
    tp = _lib.default_transport(ecut=0.521, pcut=0.01)
    n = 200000
    g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
    ref = []
    for seeds in ((1802, 9373), NULL_SEEDS):
        o = oracle_lib.OracleSimulation(BLOB521, g, src[0], tp, nstreams=256,
                                        nthreads=oracle_threads())
        o.seed(*seeds)
        o.run(n)
        ref.append(o.tallies())
        del o
    sim = Simulation(Tables(BLOB521), xb, yb, zb, med, rhor, src, tp)
    sim.run(n)
    check_parity(ref[0], sim.tallies(), sim.counters(), label="c2 full size",
                 total_rtol=0.003, null=ref[1])


@pytest.mark.gpu
def test_c1_full_size_parity(engine):
    """C1 at full size: 128^3 x 0.25 cm water / lung (0.26) / bone (1.85) /
    water slabs, 10 x 10 cm^2 parallel beam of the 60-bin falling 0.1-6 MeV
    spectrum, ECUT 0.521 MeV, PCUT 0.01 MeV, coupled transport on the
    default event schedule, 4e6 histories against the oracle.  chi^2/dof <=
    1.1 over the voxels above 10 % of Dmax; the 2-sigma fraction against
    that of a second, independent oracle run (most voxels get their dose
    from a few tens of histories at this size)."""
    xb, yb, zb, med, rhor = workloads.slab_phantom(128, 0.25)
    e, w = workloads.falling_spectrum()
    src = make_source(0, 0.0, spectrum=(e, w), x0=-5, x1=5, y0=-5, y1=5, z=-1.0)
    tp = _lib.default_transport(ecut=0.521, pcut=0.01)
    n = 4000000
    g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
    ref = []
    for seeds in ((1802, 9373), NULL_SEEDS):
        o = oracle_lib.OracleSimulation(BLOB521, g, src[0], tp, nstreams=256,
                                        nthreads=oracle_threads())
        o.seed(*seeds)
        o.run(n)
        ref.append(o.tallies())
        del o
    sim = Simulation(Tables(BLOB521), xb, yb, zb, med, rhor, src, tp)
    sim.seed(33, 97)
    sim.run(n)
    b = sim.tallies()
    assert b[2] == n
    check_parity(ref[0], b, sim.counters(), label="c1 full size", null=ref[1],
                 total_rtol=0.003)
    eb = sim.energy_balance()
    assert abs(eb["incident"] - eb["deposited"] - eb["escaped"]) < 1e-5 * eb["incident"]


@pytest.mark.gpu
def test_c3_batch256(engine):
    """C3: the 256-job random-field workload (edges 32-160 voxels of 1-5 mm,
    air / lung / water / bone with +-5 % density noise, photon spectra and
    4-20 MeV electron beams) in one batched event schedule, at 1 % of each
    job's histories, except three small jobs run at parity size; the batch
    is run twice, the second time as one graph launch.  Every job: its
    history count, energy balance (incident = deposited + escaped) and no
    overflow; the three: dose parity with the oracle (the engine's tallies
    hold twice the oracle's histories)."""
    from egsnrc_amd.engine import run_batched
    jobs = workloads.batch_jobs(256, hist_scale=0.01)
    tab = Tables(BLOB521)
    tp = _lib.default_transport(ecut=0.521, pcut=0.01, kerma=0)
    order = sorted(range(len(jobs)), key=lambda j: jobs[j]["n"])
    photons = [j for j in order if jobs[j]["src"]["charge"] == 0][:2]
    electrons = [j for j in order if jobs[j]["src"]["charge"] != 0][:1]
    picked = {photons[0]: 1000000, photons[1]: 1000000, electrons[0]: 200000}
    sims, keep = [], {}
    for k, j in enumerate(jobs):
        s = j["src"]
        src = make_source(s["charge"], s["energy"], spectrum=s["spectrum"], x0=s["x0"],
                          x1=s["x1"], y0=s["y0"], y1=s["y1"], z=s["z"])
        sims.append(Simulation(tab, j["xb"], j["yb"], j["zb"], j["med"], j["rhor"], src, tp,
                               schedule="event"))
        if k in picked:
            keep[k] = (j["xb"], j["yb"], j["zb"], j["med"], j["rhor"], src)
            j["nhist"] = picked[k]
        j["med"] = j["rhor"] = None
    nh = np.array([j["nhist"] for j in jobs], dtype=np.int64)
    print("batch: %d jobs, %d histories, %d voxels"
          % (len(jobs), nh.sum(), sum(s.nvox for s in sims)))
    run_batched(sims, nh)
    first = sims[0].graph_stats()
    # the second run of the batch is one graph launch: its graph was sized
    # from the first run's iteration count, and the device ended it
    run_batched(sims, nh)
    launches, iters, captured = sims[0].graph_stats()
    print("graph launches: first run %d (%d iterations); second run %d (%d of %d iterations)"
          % (first[0], first[1], launches, iters, captured))
    assert launches == 1 and 0 < iters <= captured
    c = sims[0].counters()
    assert c[5] == 0, "stack or bank overflow"
    # a history that touches more voxels than its slot holds adds the rest
    # straight into the sums (the mean stays exact): allowed in 1e-5 of them
    print("history tally table overflows: %d in %d histories" % (c[6], 2 * nh.sum()))
    assert c[6] <= 1e-5 * 2 * nh.sum(), "history tally table overflow"
    worst = 0.0
    for j, s in enumerate(sims):
        t = s.tallies()
        assert t[2] == 2 * nh[j]
        eb = s.energy_balance()
        assert eb["incident"] > 0 and eb["deposited"] > 0
        r = abs(eb["incident"] - eb["deposited"] - eb["escaped"]) / eb["incident"]
        worst = max(worst, r)
    print("largest relative energy-balance residual over the jobs: %.2e" % worst)
    assert worst < 1e-5
    for k, n in picked.items():
        xb, yb, zb, med, rhor, src = keep[k]
        g, gk = _lib.make_geometry(xb, yb, zb, med, rhor)
        ref = []
        for seeds in (None, NULL_SEEDS):
            o = oracle_lib.OracleSimulation(BLOB521, g, src[0], tp, nstreams=256,
                                            nthreads=oracle_threads())
            if seeds:
                o.seed(*seeds)
            o.run(n)
            ref.append(o.tallies())
            del o
        check_parity(ref[0], sims[k].tallies(), label="c3 job %d (n=%d, %s)"
                     % (k, jobs[k]["n"], "photons" if jobs[k]["src"]["charge"] == 0
                        else "electrons %.1f MeV" % jobs[k]["src"]["energy"]), null=ref[1])

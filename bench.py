"""Benchmark of the MI355X voxel transport engine.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

A step is one pass of the transport path over one batch of primary
histories, with tables, phantom and source already resident in HBM.  The
timed region is K steps between two device synchronisations.  Prints one
JSON line with histories/s, the efficiency 1/(sigma^2 T), the roofline of
the transport kernel and a CPU baseline (the FP64 oracle run on this host's
cores over a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = ("histories/s + efficiency 1/(\u03c3\u00b2\u00b7T) at 2\u03c3 dose agreement; "
          "density-grad evals/s")

WORKLOADS = {
    # heterogeneous slab phantom: 128^3 x 2.5 mm, water/lung/bone/water
    # along the beam, 10x10 cm^2 parallel photon field with a 60-bin
    # spectrum falling linearly from 0.1 to 6 MeV; coupled e-/gamma
    # transport with ECUT 0.521, PCUT 0.01 (default transport options,
    # spin effects off)
    "slab128": dict(blob="521icru.egstab", hist_per_step=2_000_000),
    # photon-only kerma phantom: 64^3 x 4 mm, density field in [0.2, 2.0],
    # 2 MeV 8x8 cm^2 parallel beam, ECUT 20 MeV (electrons deposited locally)
    "kerma64": dict(blob="521icru.egstab", hist_per_step=4_000_000),
    # C0: 1.25 MeV photon pencil beam on 30^3 x 1 cm water, ECUT 0.7,
    # PCUT 0.01, Rayleigh off (the reference's own CPU-runnable case)
    "c0": dict(blob="700icru.egstab", hist_per_step=4_000_000),
    # C2: 20 MeV electrons, 10x10 cm2, on 256^3 x 0.2 cm water with a 3 cm
    # aluminium cube at 3 cm depth; ECUT 0.521, PCUT 0.01, spin off
    "c2": dict(blob="521icru.egstab", hist_per_step=500_000),
    # many small jobs: 256 phantoms of 32-160^3 voxels of 1-5 mm (random
    # air/lung/water/bone fields), photon and 4-20 MeV electron beams,
    # histories log-uniform in [1e5, 1e7] per job scaled by --hist-scale,
    # all run as one batch (egsgpu_run_batched)
    "batch256": dict(blob="521icru.egstab", hist_per_step=0),
}


def build_workload(name):
    from egsnrc_amd import _lib
    from egsnrc_amd.engine import make_source
    from egsnrc_amd import workloads
    if name == "slab128":
        xb, yb, zb, med, rhor = workloads.slab_phantom(128, 0.25)
        e, w = workloads.falling_spectrum()
        src = make_source(0, 0.0, spectrum=(e, w), x0=-5, x1=5, y0=-5, y1=5, z=-1.0)
        tp = _lib.default_transport(ecut=0.521, pcut=0.01, kerma=0)
        cfg = {"workload": "slab128: 128^3 x 0.25 cm phantom, water/lung(0.26)/bone(1.85)/water "
                           "slabs along z, 10x10 cm2 parallel photon field, 60-bin spectrum "
                           "0.1-6 MeV falling linearly, ECUT 0.521 MeV, PCUT 0.01 MeV, coupled "
                           "photon/electron transport (PRESTA-II, exact BCA, spin off)",
               "voxels": 128 ** 3}
        return xb, yb, zb, med, rhor, src, tp, cfg
    if name == "c0":
        xb, yb, zb, med, rhor = workloads.water_cube(30, 1.0)
        src = make_source(0, 1.25, shape="point", x0=0.0, y0=0.0, z=-1.0)
        tp = _lib.default_transport(ecut=0.7, pcut=0.01, rayleigh=0)
        cfg = {"workload": "c0: 1.25 MeV photon pencil beam, 30^3 x 1 cm water, ECUT 0.7 MeV, "
                           "PCUT 0.01 MeV, Rayleigh off, bound Compton, spin off",
               "voxels": 30 ** 3}
        return xb, yb, zb, med, rhor, src, tp, cfg
    if name == "c2":
        xb, yb, zb, med, rhor = workloads.al_cube_phantom(256, 0.2)
        src = make_source(-1, 20.0, x0=-5, x1=5, y0=-5, y1=5, z=-1.0)
        tp = _lib.default_transport(ecut=0.521, pcut=0.01)
        cfg = {"workload": "c2: 20 MeV electrons 10x10 cm2 on 256^3 x 0.2 cm water with a 3 cm "
                           "Al cube at 3 cm depth, ECUT 0.521, PCUT 0.01, ESTEPE 0.25, spin off",
               "voxels": 256 ** 3}
        return xb, yb, zb, med, rhor, src, tp, cfg
    if name == "kerma64":
        xb, yb, zb, med, rhor = workloads.kerma_phantom(64, 0.4)
        src = make_source(0, 2.0, x0=-4, x1=4, y0=-4, y1=4, z=-1.0)
        tp = _lib.default_transport(ecut=20.0, kerma=1)
        cfg = {"workload": "kerma64: 64^3 x 0.4 cm water, density GRF [0.2,2.0] g/cm3, "
                           "2 MeV 8x8 cm2 parallel photon beam, ECUT 20 MeV (kerma), PCUT 0.01",
               "voxels": 64 ** 3}
        return xb, yb, zb, med, rhor, src, tp, cfg
    raise SystemExit("unknown workload " + name)


def progress(msg):
    """A progress line on stderr (the JSON line alone goes to stdout)."""
    print("[bench %.0fs] %s" % (time.perf_counter() - T_START, msg), file=sys.stderr, flush=True)


T_START = time.perf_counter()


def efficiency(sum_, sum2, n, seconds):
    """1/(sigma^2 T): sigma^2 = mean relative variance over voxels with
    D > 0.5 Dmax."""
    r = sum_ / n
    v = sum2 / n - r * r
    v = np.where(v > 0, v / (n - 1), 0.0)
    m = r > 0.5 * r.max()
    rel2 = float(np.mean(v[m] / (r[m] * r[m])))
    return 1.0 / (rel2 * seconds), rel2


def cpu_baseline(name, seconds_target=15.0):
    """The FP64 oracle on the host cores, run on a bounded sample.  Returns
    the baseline record, the sample's tallies (sum, sum2, histories) and
    those of a second oracle run of the same size with other seeds (the
    null for the dose agreement; not timed)."""
    from egsnrc_amd import _lib
    from tests import oracle_lib
    xb, yb, zb, med, rhor, src, tp, cfg = build_workload(name)
    g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
    # every core of the host, up to the 16 that one GPU's share of the box
    # provides
    cores = min(len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
                else (os.cpu_count() or 1), 16)
    blob = os.path.join(ROOT, "egsnrc_amd", "data", WORKLOADS[name]["blob"])
    o = oracle_lib.OracleSimulation(blob, g, src[0], tp, nstreams=4 * cores, nthreads=cores)
    o.seed(4711, 20000)   # independent of the engine's default streams
    n = 20000
    t = time.perf_counter()
    o.run(n)
    dt = time.perf_counter() - t
    # scale the sample to about seconds_target of CPU time
    n2 = int(min(5e7, max(n, n * seconds_target / max(dt, 1e-3))))
    t = time.perf_counter()
    o.run(n2)
    dt = time.perf_counter() - t
    rec = {"value": n2 / dt, "unit": "histories/s", "cores": cores, "kind": "port",
           "sample": "%d histories of the same workload (FP64 oracle, %d threads, %.1f s)"
                     % (n2, cores, dt),
           "note": "the oracle (a scalar FP64 restatement of the reference's routines) "
                   "stands in for the reference's CPU path, which cannot be built here: "
                   "egs++ needs the configure-generated egs_config1.h and the Mortran "
                   "core mortran3-generated Fortran; 16 threads = one GPU's share of the "
                   "host"}
    tallies = o.tallies()
    o2 = oracle_lib.OracleSimulation(blob, g, src[0], tp, nstreams=4 * cores, nthreads=cores)
    o2.seed(2718, 31415)
    o2.run(n + n2)
    return rec, tallies, o2.tallies()


def dose_agreement(gpu, ref, frac=0.1):
    """Per-voxel z of the engine's tallies against the oracle's over the
    voxels above frac of Dmax: chi^2/dof and the fraction beyond 2 sigma
    (4.55 % for agreeing normal z)."""
    def st(t):
        s_, s2, n = t
        r = s_ / n
        v = s2 / n - r * r
        return r, np.where(v > 0, v / (n - 1), 0.0)
    ra, va = st(gpu)
    rb, vb = st(ref)
    m = (np.maximum(ra, rb) > frac * max(ra.max(), rb.max())) & (va + vb > 0)
    z = (ra[m] - rb[m]) / np.sqrt(va[m] + vb[m])
    return {"chi2_per_dof": float(np.mean(z * z)), "beyond_2sigma": float(np.mean(np.abs(z) > 2)),
            "voxels": int(z.size), "histories_engine": int(gpu[2]),
            "histories_oracle": int(ref[2]),
            "total_rel_diff": float(ra.sum() / rb.sum() - 1.0),
            "definition": "z = (D_engine - D_oracle) / sqrt(s_e^2 + s_o^2) over voxels above "
                          "10% of Dmax"}


def fp64_throughput(tab, name, hps, a):
    """Histories/s of the same workload with FP64 transport arithmetic
    (tallies are FP64 in both builds): one warm-up and two timed steps."""
    from egsnrc_amd.engine import Simulation
    xb, yb, zb, med, rhor, src, tp, cfg = build_workload(name)
    tp.fp64 = 1
    sim = Simulation(tab, xb, yb, zb, med, rhor, src, tp, nstreams=a.lanes, block=a.block,
                     schedule=a.schedule)
    if a.woodcock is not None:
        sim.set_woodcock(a.woodcock)
    sim.run(hps // 4)   # egsgpu_run returns when the launch has completed
    t0 = time.perf_counter()
    for _ in range(2):
        sim.run(hps)
    T = time.perf_counter() - t0
    sim.close()
    return {"histories_per_s": 2 * hps / T, "steps": 2, "histories_per_step": hps,
            "dtype": "f64"}


def atomic_bench():
    """Rate of FP64 atomic adds on this device (egsgpu_atomic_rate): the
    roof of the tally flushes and of the density-gradient sweep."""
    from egsnrc_amd import _lib
    out = {"unit": "G adds/s"}
    for key, pattern, naddr, nops in (("scattered_64cube", 0, 64 ** 3, 1 << 28),
                                      ("scattered_128cube", 0, 128 ** 3, 1 << 28),
                                      ("coalesced_wave", 2, 1 << 24, 1 << 28),
                                      ("one_address_per_wave", 1, 1 << 20, 1 << 24)):
        rate, total, issued = _lib.atomic_rate(pattern, naddr, nops)
        if total != 2 * issued:
            raise RuntimeError("atomic_rate: %s lost adds (%g of %d)" % (key, total, 2 * issued))
        out[key] = rate / 1e9
    return out


def grad_bench(nhist, lanes=0, fd_seeds=4):
    """Density-gradient evaluations/s on the C4 problem: 64^3 x 0.4 cm
    water with densities from a smoothed random field in [0.2, 2.0], 2 MeV
    8x8 cm^2 parallel beam, kerma mode; target dose from a second field
    with a 10 % perturbation.  One evaluation = forward transport + replay
    of nhist histories (egsgpu_density_grad), with Woodcock photon tracking
    (the default here) and with voxel-by-voxel tracking.  The finite-
    difference check: the score-function directional derivatives along
    five smooth random directions against central differences of the loss
    with common random numbers, h = 3e-2 rho, averaged over fd_seeds seeds
    (relative L2 of the difference)."""
    from egsnrc_amd import _lib
    from egsnrc_amd.engine import Simulation, Tables, make_source
    from egsnrc_amd import workloads
    n, dx = 64, 0.4
    xb, yb, zb, med, rhor = workloads.kerma_phantom(n, dx, seed=4)
    _, _, _, _, rho2 = workloads.kerma_phantom(n, dx, seed=5)
    trho = np.clip(rhor * (1 + 0.1 * (rho2 - rho2.mean()) / rho2.std()), 0.2, 2.0)
    src = make_source(0, 2.0, x0=-4, x1=4, y0=-4, y1=4, z=-1.0)
    tp = _lib.default_transport(ecut=20.0, kerma=1)
    tab = Tables(os.path.join(ROOT, "egsnrc_amd", "data", "521icru.egstab"))
    vol = dx ** 3

    def sim_for(r, seed, woodcock):
        s = Simulation(tab, xb, yb, zb, med, r, src, tp, nstreams=lanes, schedule="history")
        if woodcock:
            s.set_woodcock(0.0)
        s.seed(seed, 9373)
        return s

    def dose(r, seed, nh):
        s = sim_for(r, seed, True)
        s.run(nh)
        su, _, nn = s.tallies()
        s.close()
        return su / nn / (vol * r)

    dstar = dose(trho, 77, min(nhist, 100_000_000))
    out = {"workload": "C4: 64^3 x 0.4 cm water, density field [0.2, 2.0], 2 MeV 8x8 cm2 "
                       "beam, kerma mode, L = sum (D - D*)^2 with D* from a 10% perturbed field",
           "histories": int(nhist),
           "note": "one evaluation = forward transport + replay from the saved RNG states"}
    for key, wc in (("woodcock", True), ("voxel", False)):
        sim = sim_for(rhor, 1802, wc)
        sim.density_grad(1_000_000, dstar)        # warm-up
        t0 = time.perf_counter()
        g, d, L = sim.density_grad(nhist, dstar)
        T = time.perf_counter() - t0
        sim.close()
        out[key] = {"seconds": T, "evals_per_s": 1.0 / T, "histories_per_s": 2.0 * nhist / T,
                    "loss": L}
    out["evals_per_s"] = out["woodcock"]["evals_per_s"]
    out["photon_tracking"] = "woodcock (majorant x 1.25); voxel-by-voxel beside it"
    if fd_seeds > 0:
        dirs = []
        for k in range(5):
            f = workloads.grf_density(n, 3.0, 100 + k)
            dirs.append(rhor * np.transpose((f - f.mean()) / f.std(), (2, 1, 0)).reshape(-1))
        h = 3e-2
        sf, fd = [], []
        for seed in range(1802, 1802 + 1000 * fd_seeds, 1000):
            s = sim_for(rhor, seed, True)
            g, _, _ = s.density_grad(nhist, dstar)
            s.close()
            sf.append([float(g @ dl) for dl in dirs])
            fd.append([float((np.sum((dose(rhor + h * dl, seed, nhist) - dstar) ** 2) -
                              np.sum((dose(rhor - h * dl, seed, nhist) - dstar) ** 2)) / (2 * h))
                       for dl in dirs])
            progress("gradient finite-difference check: seed %d done" % seed)
        sf, fd = np.mean(sf, 0), np.mean(fd, 0)
        out["fd_check"] = {"h": h, "seeds": fd_seeds, "directions": 5,
                           "rel_l2": float(np.linalg.norm(sf - fd) / np.linalg.norm(fd)),
                           "score_function": sf.tolist(), "finite_differences": fd.tolist()}
    return out


def _rel2(sim):
    """Mean relative variance of the mean dose over D > 0.5 Dmax."""
    su, s2, n = sim.tallies()
    r = su / n
    v = np.where(s2 / n - r * r > 0, (s2 / n - r * r) / (n - 1), 0.0)
    m = r > 0.5 * r.max()
    return float(np.mean(v[m] / (r[m] * r[m])))


def c3_compare(hist_scale, lanes=0, schedules=("event",)):
    """C3 in the batched persistent form against the plain form.

    Batched: the 256 jobs in one egsgpu_run_batched call (one graph launch;
    a first run of the same batch sizes the graph and is not timed).
    Plain: every job its own simulation, run alone, with the given
    schedules (the event schedule, the faster of the two for these jobs, by
    default; --workload batch256 --compare-plain adds the history
    schedule), each with as many transport lanes as
    suit its size (a power of two near a quarter of its histories, between
    4096 and the default), so that a small job neither waits for nor
    allocates the state of a full-device launch (simulations made outside
    the timed region).  Efficiency 1/(sigma^2 T) with sigma^2 the mean over jobs of
    the mean relative variance over D > 0.5 Dmax, so the ratio compares the
    forms at equal per-voxel sigma."""
    from egsnrc_amd import _lib
    from egsnrc_amd.engine import Simulation, Tables, make_source, run_batched
    from egsnrc_amd import workloads
    tab = Tables(os.path.join(ROOT, "egsnrc_amd", "data", "521icru.egstab"))
    tp = _lib.default_transport(ecut=0.521, pcut=0.01, kerma=0)

    def make(j, lanes, schedule):
        s = j["src"]
        src = make_source(s["charge"], s["energy"], spectrum=s["spectrum"], x0=s["x0"],
                          x1=s["x1"], y0=s["y0"], y1=s["y1"], z=s["z"])
        return Simulation(tab, j["xb"], j["yb"], j["zb"], j["med"], j["rhor"], src, tp,
                          nstreams=lanes, schedule=schedule)

    jobs = workloads.batch_jobs(256, hist_scale=hist_scale)
    nh = np.array([j["nhist"] for j in jobs], dtype=np.int64)
    sims = [make(j, lanes if k == 0 else 0, "event") for k, j in enumerate(jobs)]
    run_batched(sims, nh)                       # sizes the run's graph
    for s in sims:
        _lib.check(_lib.lib().egsgpu_reset_tallies(s.h))
    t0 = time.perf_counter()
    run_batched(sims, nh)
    t_batch = time.perf_counter() - t0
    launches, iters, captured = sims[0].graph_stats()
    rel2_batch = float(np.mean([_rel2(s) for s in sims]))
    for s in sims:
        s.close()
    del sims
    out = {"workload": "batch256: 256 jobs, 32-160^3 voxels of 1-5 mm, random "
                       "air/lung/water/bone fields, photon 0.1-6 MeV and electron 4-20 MeV "
                       "beams, histories log-uniform 1e5-1e7 x %g" % hist_scale,
           "histories": int(nh.sum()),
           "batched": {"seconds": t_batch, "sigma2": rel2_batch,
                       "efficiency": 1.0 / (rel2_batch * t_batch),
                       "histories_per_s": float(nh.sum()) / t_batch,
                       "graph_launches": launches, "graph_iterations": iters,
                       "graph_iterations_captured": captured}}
    progress("c3: batched run done (%.2f s); plain runs" % t_batch)
    for schedule in schedules:
        T, rel2 = 0.0, []
        for j, n in zip(jobs, nh):
            lanes_j = int(min(262144, max(4096, 2 ** round(np.log2(max(1, n) / 4)))))
            sim = make(j, lanes_j, schedule)
            sim.run(min(int(n), 1000))          # allocation and first launch
            _lib.check(_lib.lib().egsgpu_reset_tallies(sim.h))
            t0 = time.perf_counter()
            sim.run(int(n))
            T += time.perf_counter() - t0
            rel2.append(_rel2(sim))
            sim.close()
        rel2 = float(np.mean(rel2))
        out["plain_" + schedule] = {"seconds": T, "sigma2": rel2, "efficiency": 1.0 / (rel2 * T),
                                    "histories_per_s": float(nh.sum()) / T}
        progress("c3: plain %s schedule done (%.2f s)" % (schedule, T))
    best = max(out["plain_" + sc]["efficiency"] for sc in schedules)
    out["batched_over_best_plain_efficiency"] = out["batched"]["efficiency"] / best
    out["note"] = ("plain = each job alone in its own simulation, one after the other, "
                   "with lanes sized to the job; the best of the plain schedules run (%s) is "
                   "the reference" % ", ".join(schedules))
    return out


def bench_batch(a):
    """The batch workload: all jobs advanced by one egsgpu_run_batched call
    per step."""
    from egsnrc_amd import _lib
    from egsnrc_amd.engine import Simulation, Tables, make_source, run_batched
    from egsnrc_amd import workloads
    tab = Tables(os.path.join(ROOT, "egsnrc_amd", "data", "521icru.egstab"))
    jobs = workloads.batch_jobs(256, hist_scale=a.hist_scale)
    tp = _lib.default_transport(ecut=0.521, pcut=0.01, kerma=0)
    sims = []
    for k, j in enumerate(jobs):
        s = j["src"]
        src = make_source(s["charge"], s["energy"], spectrum=s["spectrum"], x0=s["x0"],
                          x1=s["x1"], y0=s["y0"], y1=s["y1"], z=s["z"])
        sims.append(Simulation(tab, j["xb"], j["yb"], j["zb"], j["med"], j["rhor"], src, tp,
                               nstreams=a.lanes if k == 0 else 0, schedule="event"))
        j["med"] = j["rhor"] = None
    nh = np.array([j["nhist"] for j in jobs], dtype=np.int64)
    for _ in range(a.warmup):
        run_batched(sims, nh)
    for s in sims:
        _lib.check(_lib.lib().egsgpu_reset_tallies(s.h))
    kms = np.zeros(4)
    klaunch = np.zeros(4, dtype=np.int64)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        run_batched(sims, nh)
        for k in range(4):
            ms, nl = sims[0].kernel_time(k)
            kms[k] += ms
            klaunch[k] += nl
    T = time.perf_counter() - t0
    total = int(nh.sum()) * a.steps
    rel2 = []
    for s in sims:
        su, s2, n = s.tallies()
        r = su / n
        v = np.where(s2 / n - r * r > 0, (s2 / n - r * r) / (n - 1), 0.0)
        m = r > 0.5 * r.max()
        rel2.append(float(np.mean(v[m] / (r[m] * r[m]))))
    rel2 = float(np.mean(rel2))
    knames = ["k_transport<float>", "k_ev_source<float>", "k_ev_photon<float>",
              "k_ev_electron<float>"]
    call = sims[0].counters()
    if klaunch.sum() > 0:
        kdom = int(np.argmax(kms))
        dc = sims[0].kernel_counters(kdom)
        kname, kt, nl = knames[kdom], kms[kdom] / 1e3, int(klaunch[kdom])
    else:
        # the batch ran from a HIP graph, which has no per-kernel times:
        # all kernels' events over the whole timed region
        dc = call
        kname, kt, nl = "all kernels (HIP graph)", T, 0
    alg_bytes = 5.0 * dc[1] + 16.0 * dc[2] + 32.0 * dc[3]
    achieved = alg_bytes / kt / 1e9
    out = {
        "metric": METRIC, "value": total / T, "unit": "histories/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * T / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "tally_dtype": "f64", "data": "synthetic phantoms and sources",
        "config": {"workload": "batch256: 256 jobs, 32-160^3 voxels of 1-5 mm, random "
                               "air/lung/water/bone fields, photon 0.1-6 MeV and electron "
                               "4-20 MeV beams, histories log-uniform 1e5-1e7 x %g per step"
                               % a.hist_scale,
                   "jobs": len(jobs), "voxels": int(sum(j["n"] ** 3 for j in jobs)),
                   "histories_per_step": int(nh.sum()), "parallelism": "none",
                   "schedule": "event"},
        "efficiency": {"value": 1.0 / (rel2 * T), "unit": "1/(sigma^2 s)", "sigma2": rel2,
                       "definition": "mean over jobs of the mean relative variance over "
                                     "D > 0.5 Dmax"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": kname, "kernel_launches": nl,
                     "kernel_ms_avg": 1e3 * kt / max(1, nl)},
        "kernels_ms_per_step": {knames[k]: kms[k] / a.steps for k in range(4) if klaunch[k]},
        "counters": {"photon_steps": int(call[0]), "voxel_crossings": int(call[1]),
                     "deposits": int(call[2]), "flushes": int(call[3]),
                     "electron_steps": int(call[4]), "overflows": int(call[5]),
                     "tally_overflows": int(call[6])},
    }
    if a.compare_plain:
        out["c3_compare"] = c3_compare(a.hist_scale, a.lanes, ("event", "history"))
    print(json.dumps(out))


def _same_kernel(a, b):
    """Kernel names that agree in base name and first template argument
    (the bench names k_ev_electron<float>, the profiler
    k_ev_electron<float, 2, false>)."""
    def parts(n):
        n = _kernel_key(n)
        base, _, rest = n.partition("<")
        return base, rest.split(",")[0].rstrip(">").strip()
    return parts(a) == parts(b)


def _kernel_key(name):
    """Kernel name without return type, namespace and argument list."""
    return name.split("(")[0].replace("void ", "").replace("egsgpu::", "").strip()


def _pmc_per_dispatch(d, counter):
    """Per-kernel (dispatches, mean counter value) from one rocprofv3 --pmc
    output directory written with `-o run --output-format csv`.  Dispatches
    whose value is below 1e-4 of the kernel's largest are left out: they are
    the iterations of a run's graph that the device found already finished,
    whose kernels return at once."""
    import csv
    per = {}
    with open(os.path.join(d, "run_counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            k = _kernel_key(r["Kernel_Name"])
            per.setdefault(k, {})
            per[k][r["Dispatch_Id"]] = per[k].get(r["Dispatch_Id"], 0.0) + float(r["Counter_Value"])
    out = {}
    for k, dv in per.items():
        mx = max(dv.values())
        v = [x for x in dv.values() if x >= 1e-4 * mx] or [0.0]
        out[k] = (len(v), sum(v) / len(v))
    return out


def pmc_summary(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, command):
    """Markdown summary of HBM traffic per dispatch.

    FETCH_SIZE and WRITE_SIZE (KiB per dispatch, one rocprofv3 pass each)
    are scaled by what the same counters read for the device-to-device
    copy of `--copy-calibration` (3 copies of 1 GiB: 3 GiB read, 3 GiB
    written).  The line `traffic_bytes_per_launch <kernel>: <bytes>` is
    what the bench reads back as roofline.traffic."""
    cf = _pmc_per_dispatch(copy_fetch_dir, "FETCH_SIZE")
    cw = _pmc_per_dispatch(copy_write_dir, "WRITE_SIZE")
    ck = [k for k in cf if "copyBuffer" in k][0]
    copy_bytes = float(1 << 30)
    f_fac = copy_bytes / (cf[ck][1] * 1024.0)
    w_fac = copy_bytes / (cw[ck][1] * 1024.0)
    fe = _pmc_per_dispatch(fetch_dir, "FETCH_SIZE")
    wr = _pmc_per_dispatch(write_dir, "WRITE_SIZE")
    out = ["Commands (from the repository root, `TMPDIR=/tmp`):", ""]
    for c, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        out.append("    rocprofv3 --pmc %s -d %s -o run --output-format csv -- %s"
                   % (c, d, command))
    for c, d in (("FETCH_SIZE", copy_fetch_dir), ("WRITE_SIZE", copy_write_dir)):
        out.append("    rocprofv3 --pmc %s -d %s -o run --output-format csv -- "
                   "python3 bench.py --copy-calibration" % (c, d))
    out += ["", "Copy calibration (`%s`, 1 GiB read + 1 GiB written per dispatch): "
            "FETCH_SIZE reads %.1f KiB, WRITE_SIZE %.1f KiB per dispatch; scale factors "
            "FETCH x%.3f, WRITE x%.3f." % (ck, cf[ck][1], cw[ck][1], f_fac, w_fac), "",
            "| kernel | dispatches | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch "
            "| HBM read B/dispatch | HBM written B/dispatch | HBM total B/dispatch |",
            "|---|---|---|---|---|---|---|"]
    lines = []
    for k in sorted(fe, key=lambda k: -fe[k][1]):
        if k not in wr:
            continue
        rb = fe[k][1] * 1024.0 * f_fac
        wb = wr[k][1] * 1024.0 * w_fac
        out.append("| `%s` | %d | %.4g | %.4g | %.4g | %.4g | %.4g |"
                   % (k, fe[k][0], fe[k][1], wr[k][1], rb, wb, rb + wb))
        if k.startswith("k_"):
            lines.append("traffic_bytes_per_launch %s: %.6g" % (k, rb + wb))
    # the algorithmic bytes per launch of the dominant kernel in the profiled
    # run (its bench line, printed into <fetch_dir>.log), so that a run with
    # another launch shape can scale the traffic to its own launches
    log = fetch_dir.rstrip("/") + ".log"
    if os.path.exists(log):
        bl = [l for l in open(log) if l.startswith("{")]
        if bl:
            rf = json.loads(bl[-1])["roofline"]
            lines.append("algorithmic_bytes_per_launch %s: %.6g"
                         % (rf["kernel"], rf["algorithmic_bytes_per_launch"]))
    return "\n".join(out + [""] + lines) + "\n"


SQ_COUNTERS = ("SQ_WAVES SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_WAIT_ANY SQ_ACTIVE_INST_VALU "
               "SQ_THREAD_CYCLES_VALU SQ_INSTS_VALU GRBM_GUI_ACTIVE")


def sq_summary(sq_dir, command):
    """Markdown summary of one rocprofv3 --pmc pass over SQ_COUNTERS: per
    kernel, the VALU lane utilisation SQ_THREAD_CYCLES_VALU / (64
    SQ_ACTIVE_INST_VALU) (the share of the 64 lanes doing work when a VALU
    instruction issues) and the share of wave cycles spent waiting
    (SQ_WAIT_ANY / SQ_WAVE_CYCLES).  The line `valu_lane_utilisation
    <kernel>: <fraction>` is what the bench reads back."""
    import csv
    agg, disp = {}, {}
    with open(os.path.join(sq_dir, "run_counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            k = _kernel_key(r["Kernel_Name"])
            agg.setdefault(k, {})
            agg[k][r["Counter_Name"]] = agg[k].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
            disp.setdefault(k, set()).add(r["Dispatch_Id"])
    out = ["Command (from the repository root, `TMPDIR=/tmp`):", "",
           "    rocprofv3 --pmc %s -d %s -o run --output-format csv -- %s"
           % (SQ_COUNTERS, sq_dir, command), "",
           "| kernel | dispatches | waves/dispatch | VALU instructions/dispatch | VALU lane "
           "utilisation | waiting share of wave cycles |", "|---|---|---|---|---|---|"]
    lines = []
    for k in sorted(agg, key=lambda k: -agg[k].get("SQ_WAVE_CYCLES", 0)):
        d, n = agg[k], len(disp[k])
        if d.get("SQ_ACTIVE_INST_VALU", 0) <= 0 or not k.startswith("k_"):
            continue
        util = d["SQ_THREAD_CYCLES_VALU"] / (64.0 * d["SQ_ACTIVE_INST_VALU"])
        wait = d["SQ_WAIT_ANY"] / d["SQ_WAVE_CYCLES"]
        out.append("| `%s` | %d | %.0f | %.4g | %.3f | %.3f |"
                   % (k, n, d["SQ_WAVES"] / n, d["SQ_INSTS_VALU"] / n, util, wait))
        lines.append("valu_lane_utilisation %s: %.4f" % (k, util))
        lines.append("wait_share %s: %.4f" % (k, wait))
    return "\n".join(out + [""] + lines) + "\n"


def read_profile_value(kind, kernel, workload, stem):
    """A `<kind> <kernel>: <value>` line written by sq_summary or
    pmc_summary under profiles/<stem>_<workload>.md (kernel matched by name
    and precision), with the file it came from."""
    import re
    p = os.path.join(ROOT, "profiles", "%s_%s.md" % (stem, workload))
    if not os.path.exists(p):
        return None, None
    for l in open(p):
        m = re.match(r"%s (.+): (\S+)\s*$" % kind, l)
        if m and _same_kernel(m.group(1), kernel):
            return float(m.group(2)), os.path.relpath(p, ROOT)
    return None, None


def trace_summary(trace_dir, bench_log, command):
    """Markdown summary of a `rocprofv3 --kernel-trace --stats` run
    (`-o run --output-format csv`) next to the bench line printed by the
    same run, comparing the trace's average duration of the dominant kernel
    with the HIP-event average the bench measured."""
    import csv
    rows = list(csv.DictReader(open(os.path.join(trace_dir, "run_kernel_stats.csv"))))
    line = json.loads([l for l in open(bench_log) if l.startswith("{")][-1])
    rf = line["roofline"]
    # the dominant kernel's launches that did work (the iterations of a run's
    # graph that the device found finished return within microseconds)
    work = []
    for r in csv.DictReader(open(os.path.join(trace_dir, "run_kernel_trace.csv"))):
        if _same_kernel(r["Kernel_Name"], rf["kernel"]):
            dt = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
            if dt > 0.05:
                work.append(dt)
    out = ["Command (from the repository root, `TMPDIR=/tmp`):", "",
           "    rocprofv3 --kernel-trace --stats -d %s -o run --output-format csv -- %s"
           % (trace_dir, command), "",
           "| kernel | calls | total (us) | average (us) | min (us) | max (us) | % |",
           "|---|---|---|---|---|---|---|"]
    dom = None
    for r in rows:
        k = _kernel_key(r["Name"])
        out.append("| `%s` | %s | %.1f | %.1f | %.1f | %.1f | %.3f |"
                   % (k, r["Calls"], float(r["TotalDurationNs"]) / 1e3,
                      float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3,
                      float(r["MaxNs"]) / 1e3, float(r["Percentage"])))
        if _same_kernel(k, rf["kernel"]):
            dom = float(r["AverageNs"]) / 1e6
    out += ["", "Bench line of the same run: %.0f histories/s, %.1f ms per step; "
            "dominant kernel `%s`, %d launches, HIP-event average %.3f ms%s%s; "
            "roofline achieved %.1f GB/s (%.2f %% of %.0f GB/s)."
            % (line["value"], line["ms_per_step"], rf["kernel"], rf["kernel_launches"],
               rf["kernel_ms_avg"],
               "" if dom is None else ", trace average over all launches %.3f ms" % dom,
               "" if not work else "; trace average over the %d launches that did work "
               "(longer than 50 us) %.3f ms (HIP-event / trace = %.3f)"
               % (len(work), sum(work) / len(work), rf["kernel_ms_avg"] / (sum(work) / len(work))),
               rf["achieved"], 100 * rf["frac"], rf["peak"])]
    return "\n".join(out) + "\n"


def read_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` (matched by name and precision)
    recorded by pmc_summary for this workload under profiles/, the
    algorithmic bytes per launch of the profiled run, and the file they
    came from; Nones if there is none."""
    import re
    p = os.path.join(ROOT, "profiles", "hbm_traffic_%s.md" % workload)
    if not os.path.exists(p):
        return None, None, None
    traffic = alg = None
    for l in open(p):
        m = re.match(r"(traffic|algorithmic)_bytes_per_launch (.+): (\S+)\s*$", l)
        if m and _same_kernel(m.group(2), kernel):
            if m.group(1) == "traffic":
                traffic = float(m.group(3))
            else:
                alg = float(m.group(3))
    if traffic is None:
        return None, None, None
    return traffic, alg, os.path.relpath(p, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="slab128")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lanes", type=int, default=0,
                    help="transport lanes (ranmar streams); 0 = library default")
    ap.add_argument("--block", type=int, default=0, help="workgroup size; 0 = default")
    ap.add_argument("--hist", type=int, default=0, help="histories per step; 0 = workload default")
    ap.add_argument("--schedule", choices=["history", "event", "history-plain"], default=None,
                    help="transport schedule; default = the library's choice")
    ap.add_argument("--waves", default="",
                    help="event kernels' resident waves per SIMD as P,E (2-4 each)")
    ap.add_argument("--slots", type=int, default=0, help="event schedule history slots")
    ap.add_argument("--graph", action="store_true",
                    help="time the event schedule in HIP-graph mode (no per-kernel times); "
                         "by default the timed steps launch kernels from the host with HIP "
                         "events around each launch, and a short graph-mode pass is "
                         "reported beside them")
    ap.add_argument("--batching", default="",
                    help="event kernels: photon steps per turn, interaction batch, max wait, "
                         "defer below")
    ap.add_argument("--autotune", type=int, default=-1,
                    help="run the event-schedule autotuner (egsgpu_autotune: register budget, "
                         "queue capacity, workgroup size, deferral threshold, interaction batch, "
                         "photon steps) with this many probe histories before the warm-up, "
                         "outside the timed region; -1 = one step's histories, 0 = library "
                         "defaults")
    ap.add_argument("--compare-plain", action="store_true",
                    help="batch256: also run the jobs one at a time with the plain kernel")
    ap.add_argument("--grad-hist", type=int, default=100_000_000,
                    help="histories of the density-gradient evaluation reported beside "
                         "the main line (C4 problem); 0 = skip")
    ap.add_argument("--grad-fd-seeds", type=int, default=4,
                    help="seeds of the finite-difference check of the gradient; 0 = skip")
    ap.add_argument("--c3-scale", type=float, default=0.01,
                    help="slab128: also run the C3 batch (256 jobs) at this fraction of its "
                         "histories in the batched form and job by job; 0 = skip")
    ap.add_argument("--hist-scale", type=float, default=0.01,
                    help="batch256: fraction of each job's histories run per step")
    ap.add_argument("--precision-report", action="store_true",
                    help="only run the workload with FP32 and FP64 transport (same seeds) "
                         "and print the per-voxel difference summary")
    ap.add_argument("--woodcock", type=float, default=None,
                    help="photon flights by Woodcock tracking with this majorant factor "
                         "(0 = library default 1.25); also used by the gradient evaluation")
    ap.add_argument("--howfarless", action="store_true",
                    help="DOSXYZnrc HOWFARLESS transport (homogeneous workloads: c0)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device three times (the reference "
                         "kernel for scaling HBM traffic counters)")
    ap.add_argument("--pmc-summary", nargs=4, default=None,
                    metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR"),
                    help="only print the HBM traffic summary of four rocprofv3 --pmc runs")
    ap.add_argument("--sq-summary", default=None, metavar="SQ_DIR",
                    help="only print the VALU summary of a rocprofv3 --pmc run over the SQ "
                         "counters")
    ap.add_argument("--trace-summary", nargs=2, default=None, metavar=("TRACE_DIR", "BENCH_LOG"),
                    help="only print the summary of a rocprofv3 --kernel-trace --stats run")
    ap.add_argument("--profiled-command", default="python3 bench.py",
                    help="command line quoted in --pmc-summary / --trace-summary output")
    a = ap.parse_args()
    if a.pmc_summary:
        sys.stdout.write(pmc_summary(*a.pmc_summary, a.profiled_command))
        return
    if a.trace_summary:
        sys.stdout.write(trace_summary(*a.trace_summary, a.profiled_command))
        return
    if a.sq_summary:
        sys.stdout.write(sq_summary(a.sq_summary, a.profiled_command))
        return
    if a.copy_calibration:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 << 28, "copy_bytes_written": 4 << 28,
                          "copies": 3}))
        return
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(histories would shard across GPUs with one all-reduce of the tallies)",
              file=sys.stderr)
        sys.exit(2)
    from egsnrc_amd import _lib
    from egsnrc_amd.engine import Simulation, Tables
    name = a.workload
    if name == "batch256":
        return bench_batch(a)
    xb, yb, zb, med, rhor, src, tp, cfg = build_workload(name)
    W = WORKLOADS[name]
    tab = Tables(os.path.join(ROOT, "egsnrc_amd", "data", W["blob"]))
    if a.precision_report:
        from egsnrc_amd.precision_report import precision_report
        n = a.hist or 1_000_000
        t0 = time.perf_counter()
        rep, _ = precision_report(tab, xb, yb, zb, med, rhor, src, tp, n)
        rep.update({"workload": cfg["workload"], "seconds": time.perf_counter() - t0,
                    "comparison": "FP32 vs FP64 transport, same seeds and streams, "
                                  "voxels above 10% of Dmax; z = diff / combined sigma"})
        print(json.dumps(rep))
        return
    if a.howfarless:
        tp.howfarless = 1
        a.schedule = None
    sim = Simulation(tab, xb, yb, zb, med, rhor, src, tp, nstreams=a.lanes, block=a.block,
                     schedule=a.schedule)
    hps = a.hist or W["hist_per_step"]
    if a.waves or a.slots:
        wp, we = (int(x) for x in a.waves.split(",")) if a.waves else (0, 0)
        _lib.check(_lib.lib().egsgpu_set_event_shape(sim.h, wp, we, a.slots))
    if a.batching:
        ks, bt, mw, db = (int(x) for x in a.batching.split(","))
        _lib.check(_lib.lib().egsgpu_set_event_batching(sim.h, ks, bt, mw, db))
    if not a.graph:
        _lib.check(_lib.lib().egsgpu_set_event_graph(sim.h, 0))
    if a.woodcock is not None:
        sim.set_woodcock(a.woodcock)
    tuned = sim.autotune(hps if a.autotune < 0 else a.autotune) if a.autotune else None
    for _ in range(a.warmup):
        sim.run(hps)
    progress("warm-up done; timing %d steps" % a.steps)
    _lib.check(_lib.lib().egsgpu_reset_tallies(sim.h))
    kms = np.zeros(4)
    klaunch = np.zeros(4, dtype=np.int64)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        sim.run(hps)
        for k in range(4):
            ms, nl = sim.kernel_time(k)
            kms[k] += ms
            klaunch[k] += nl
    t1 = time.perf_counter()
    T = t1 - t0
    s, s2, n = sim.tallies()
    hist_s = a.steps * hps / T
    eff, rel2 = efficiency(s, s2, n, T)
    # the dominant kernel (the tallies, and so the counters, were reset
    # after the warm-up)
    kdom = int(np.argmax(kms))
    dc = sim.kernel_counters(kdom)
    call = sim.counters()
    # algorithmic HBM bytes: 5 B per voxel crossing, 16 B per deposit,
    # 32 B per (history, voxel) flush
    alg_bytes = 5.0 * dc[1] + 16.0 * dc[2] + 32.0 * dc[3]
    kernel_s = kms[kdom] / 1e3 if kms[kdom] > 0 else T
    achieved = alg_bytes / kernel_s / 1e9
    knames = ["k_transport<float>", "k_ev_source<float>", "k_ev_photon<float>",
              "k_ev_electron<float>"]
    if tp.kerma and a.schedule in (None, "history") and not a.howfarless:
        knames[0] = "k_kerma<float>"   # the kerma-mode history kernel
    nl = max(1, int(klaunch[kdom]))
    traffic, traffic_alg, traffic_src = read_traffic(knames[kdom], name)
    # the profile's traffic per launch, scaled to this run's launches by the
    # algorithmic bytes (the autotuner may choose another launch shape than
    # the profiled run's)
    traffic_ratio = traffic / traffic_alg if traffic and traffic_alg else None
    if traffic_ratio:
        traffic = traffic_ratio * alg_bytes / nl
    out = {
        "metric": METRIC,
        "value": hist_s,
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * T / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "tally_dtype": "f64",
        "data": "synthetic phantom and source",
        "config": dict(cfg, histories_per_step=hps, parallelism="none",
                       schedule=a.schedule or ("history (HOWFARLESS)" if a.howfarless
                                               else "default"), graph=a.graph,
                       **({"autotuned": dict(zip(["waves_photon", "waves_electron",
                                                  "defer_below", "batch", "photon_steps",
                                                  "slots_per_stream", "workgroup"],
                                                 tuned))} if tuned else {})),
        "efficiency": {"value": eff, "unit": "1/(sigma^2 s)",
                       "sigma2": rel2, "definition": "mean relative variance over D > 0.5 Dmax"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "traffic_over_algorithmic": traffic_ratio,
                     "algorithmic_bytes_per_launch": alg_bytes / nl,
                     "kernel": knames[kdom],
                     "kernel_launches": int(klaunch[kdom]),
                     "kernel_ms_avg": kms[kdom] / max(1, klaunch[kdom]),
                     "kernel_counters": {"voxel_crossings": int(dc[1]), "deposits": int(dc[2]),
                                         "flushes": int(dc[3])}},
        "kernels_ms_per_step": {knames[k]: kms[k] / a.steps for k in range(4) if klaunch[k]},
        "counters": {"photon_steps": int(call[0]), "voxel_crossings": int(call[1]),
                     "deposits": int(call[2]), "flushes": int(call[3]),
                     "electron_steps": int(call[4]), "overflows": int(call[5]),
                     "tally_overflows": int(call[6])},
    }
    if not a.graph and klaunch[1:].sum() > 0:   # event schedule: graph-mode throughput
        _lib.check(_lib.lib().egsgpu_set_event_graph(sim.h, 1))
        t0 = time.perf_counter()
        for _ in range(2):
            sim.run(hps)
        out["graph_mode"] = {"histories_per_s": 2 * hps / (time.perf_counter() - t0),
                             "steps": 2, "note": "same steps replayed from a HIP graph"}
    if a.woodcock is not None:
        out["config"]["photon_tracking"] = "woodcock (majorant x %s)" % (a.woodcock or 1.25)
    progress("timed steps done: %.0f histories/s" % hist_s)
    if a.grad_hist and name in ("slab128", "kerma64"):
        sim.close()   # the gradient run gets the device to itself
        progress("density gradient")
        out["density_grad"] = grad_bench(a.grad_hist, a.lanes, a.grad_fd_seeds)
    if name in ("slab128", "kerma64") and not a.howfarless:
        progress("FP64 transport")
        out["fp64_transport"] = fp64_throughput(tab, name, hps, a)
        out["fp64_atomics"] = atomic_bench()
    if name == "slab128" and a.c3_scale > 0:
        progress("C3 batched vs plain")
        out["c3_batched_vs_plain"] = c3_compare(a.c3_scale, a.lanes)
    if "fp64_atomics" in out:
        # FP64 atomic adds of the dominant kernel into the dose tallies (sum
        # and sum of squares at every history-voxel flush) against the rate
        # of scattered FP64 atomic adds over a grid of the same size
        peak = out["fp64_atomics"]["scattered_128cube"]
        ach = 2.0 * dc[3] / kernel_s / 1e9
        out["atomic_roofline"] = {"achieved": ach, "peak": peak, "unit": "G adds/s",
                                  "frac": ach / peak,
                                  "kernel": knames[kdom],
                                  "what": "FP64 atomicAdd into the dose tallies, 2 per "
                                          "(history, voxel) flush; peak = measured rate of "
                                          "scattered FP64 atomic adds over 128^3 doubles"}
    valu, vsrc = read_profile_value("valu_lane_utilisation", knames[kdom], name, "valu")
    wait, _ = read_profile_value("wait_share", knames[kdom], name, "valu")
    if valu is not None:
        out["valu"] = {"lane_utilisation": valu, "wait_share": wait, "kernel": knames[kdom],
                       "source": vsrc,
                       "definition": "SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU); "
                                     "SQ_WAIT_ANY / SQ_WAVE_CYCLES"}
    if not a.no_cpu:
        progress("CPU baseline")
        out["cpu_baseline"], ref, ref2 = cpu_baseline(name)
        agree = dose_agreement((s, s2, n), ref)
        # the same statistics for two oracle runs: what agreement looks like
        # for these voxels (their errors are heavy-tailed in the low-dose
        # regions at a few million histories)
        null = dose_agreement(ref2, ref)
        agree["null_oracle_vs_oracle"] = {"chi2_per_dof": null["chi2_per_dof"],
                                          "beyond_2sigma": null["beyond_2sigma"],
                                          "voxels": null["voxels"]}
        out["dose_agreement_vs_oracle"] = agree
    print(json.dumps(out))


if __name__ == "__main__":
    main()

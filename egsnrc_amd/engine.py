"""Host-side Python mirror of the engine's C ABI (include/egsgpu.h).

``Simulation`` plays the role of EGS_AdvancedApplication for the voxel
dose path: initSimulation (egsgpu_create), runSimulation (egsgpu_run) and
the dose-scoring results (EGS_DoseScoring, egs_dose_scoring.cpp:416-479).
Every call goes to the HIP library; nothing here computes transport.
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import EgsSource, check, ptr

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


class Tables:
    """A loaded cross-section table blob (egsgpu_tables_load)."""

    def __init__(self, path):
        self.path = path
        self.h = C.c_void_p()
        check(_lib.lib().egsgpu_tables_load(path.encode(), C.byref(self.h)))

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            _lib.lib().egsgpu_tables_free(self.h)
            self.h = C.c_void_p()


def make_source(charge=0, energy=1.0, spectrum=None, shape="rectangle",
                x0=-5.0, x1=5.0, y0=-5.0, y1=5.0, z=0.0, u=(0.0, 0.0, 1.0)):
    """EGS_ParallelBeam-like source.  spectrum = (edges, weights) for a
    histogram spectrum (per-MeV weights), or None for monoenergetic."""
    keep = []
    if spectrum is None:
        s = EgsSource(charge, 0, energy, 0, None, None, 0, x0, x1, y0, y1, z, *u)
    else:
        e = np.ascontiguousarray(spectrum[0], dtype=np.float64)
        w = np.ascontiguousarray(spectrum[1], dtype=np.float64)
        keep += [e, w]
        s = EgsSource(charge, 1, 0.0, len(w), ptr(e).value, ptr(w).value, 0,
                      x0, x1, y0, y1, z, *u)
    s.shape = 0 if shape == "point" else 1
    return s, keep


def make_fano_source(charge, energy, box, rho_max, spectrum=None):
    """egs_fano_source: particles start uniformly per unit mass in
    box = (x0, x1, y0, y1, z0, z1), with isotropic directions; rho_max (g/cm^3)
    must be at least the largest voxel mass density in the box."""
    s, keep = make_source(charge, energy, spectrum=spectrum, x0=box[0], x1=box[1],
                          y0=box[2], y1=box[3], z=box[4])
    s.kind = _lib.SOURCE_FANO
    s.z1 = box[5]
    s.rho_max = rho_max
    return s, keep


def make_phsp_source(x0=0.0, y0=0.0, z=0.0):
    """A phase-space source whose particles are given by Simulation.set_phsp
    (egs_phsp_source), its plane translated by (x0, y0, z)."""
    s = EgsSource(0, 0, 0.0, 0, None, None, 0, x0, x0, y0, y0, z, 0.0, 0.0, 1.0)
    s.kind = _lib.SOURCE_PHSP
    return s, []


class Simulation:
    def __init__(self, tables, xb, yb, zb, medium, rhor, source, transport,
                 nstreams=0, block=0, schedule=None):
        self.tables = tables
        self.geom, self._gk = _lib.make_geometry(xb, yb, zb, medium, rhor)
        self.src, self._sk = source
        self.tp = transport
        self.nvox = self.geom.nx * self.geom.ny * self.geom.nz
        self.h = C.c_void_p()
        L = _lib.lib()
        check(L.egsgpu_create(tables.h, C.byref(self.geom), C.byref(self.src),
                              C.byref(self.tp), C.byref(self.h)))
        self._gk = None   # the device holds its own copy of the geometry
        if nstreams or block:
            check(L.egsgpu_set_streams(self.h, nstreams, block))
        if schedule is not None:
            check(L.egsgpu_set_schedule(self.h, {"history": _lib.SCHED_HISTORY,
                                                 "event": _lib.SCHED_EVENT,
                                                 "history-plain": _lib.SCHED_HISTORY_PLAIN}
                                        [schedule]))

    def seed(self, ixx, jxx):
        check(_lib.lib().egsgpu_seed(self.h, ixx, jxx))

    def set_phsp(self, x, y, u, v, e, wt, q):
        """Phase-space particles (kinetic energy e, weight |wt| with the sign
        of w, charge q); history i takes particle i mod n."""
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (x, y, u, v, e, wt)]
        qq = np.ascontiguousarray(q, dtype=np.int8)
        n = len(qq)
        assert all(len(a) == n for a in arrs)
        check(_lib.lib().egsgpu_set_phsp(self.h, n, *[ptr(a) for a in arrs], ptr(qq)))

    def set_woodcock(self, factor=0.0):
        """Photon flights by Woodcock tracking (egsgpu_set_photon_tracking)."""
        check(_lib.lib().egsgpu_set_photon_tracking(self.h, 1, float(factor)))

    def set_ranlux(self, luxury, seed):
        """Drive the transport with ranlux streams: stream s is
        init_ranlux(luxury, seed + s)."""
        check(_lib.lib().egsgpu_set_rng(self.h, _lib.RNG_RANLUX, luxury, seed))

    def run(self, nhist):
        check(_lib.lib().egsgpu_run(self.h, int(nhist)))

    def tallies(self):
        s = np.zeros(self.nvox)
        s2 = np.zeros(self.nvox)
        n = C.c_int64(0)
        check(_lib.lib().egsgpu_get_tallies(self.h, ptr(s), ptr(s2), C.byref(n)))
        return s, s2, n.value

    def density_grad(self, nhist, target):
        """Score-function gradient of sum_v (D_v - target_v)^2 w.r.t. the
        relative voxel densities (forward + replay on the device)."""
        target = np.ascontiguousarray(target, dtype=np.float64)
        g = np.zeros(self.nvox)
        d = np.zeros(self.nvox)
        loss = C.c_double(0)
        check(_lib.lib().egsgpu_density_grad(self.h, int(nhist), ptr(target), ptr(g), ptr(d),
                                             C.byref(loss)))
        return g, d, loss.value

    def counters(self):
        c = np.zeros(8, dtype=np.int64)
        check(_lib.lib().egsgpu_get_counters(self.h, ptr(c)))
        return c

    def energy_balance(self):
        """(incident, escaped, deposited) energy per history (MeV) and the
        standard errors of incident and escaped (egsgpu_get_energy_balance;
        deposited = the sum of the dose regions)."""
        b = np.zeros(4)
        check(_lib.lib().egsgpu_get_energy_balance(self.h, ptr(b)))
        s, s2, n = self.tallies()
        n = max(n, 1)

        def err(t, t2):
            return np.sqrt(max(t2 / n - (t / n) ** 2, 0.0) / max(n - 1, 1))
        return dict(incident=b[0] / n, escaped=b[2] / n, deposited=float(s.sum()) / n,
                    incident_err=err(b[0], b[1]), escaped_err=err(b[2], b[3]))

    def kernel_ms(self):
        ms = C.c_double(0)
        n = C.c_int32(0)
        check(_lib.lib().egsgpu_last_kernel_ms(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def autotune(self, nhist_probe=200000):
        """egsgpu_autotune; returns the chosen (waves photon, waves electron,
        defer_below, batch, photon steps, slots per stream, workgroup
        size)."""
        c = np.zeros(7, dtype=np.int32)
        check(_lib.lib().egsgpu_autotune(self.h, int(nhist_probe), ptr(c)))
        return tuple(int(x) for x in c)

    def checkpoint(self):
        """The run state (tallies, histories, RNG streams) as bytes."""
        n = C.c_size_t(0)
        check(_lib.lib().egsgpu_checkpoint(self.h, None, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(_lib.lib().egsgpu_checkpoint(self.h, buf, C.byref(n)))
        return buf.raw[:n.value]

    def restore(self, data):
        buf = C.create_string_buffer(data, len(data))
        check(_lib.lib().egsgpu_restore(self.h, buf, len(data)))

    def kernel_counters(self, kernel):
        c = np.zeros(8, dtype=np.int64)
        check(_lib.lib().egsgpu_kernel_counters(self.h, kernel, ptr(c)))
        return c

    def kernel_time(self, kernel):
        """(ms, launches) of one kernel in the last run: 0 history transport,
        1 event source, 2 event photon, 3 event charged particle."""
        ms = C.c_double(0)
        n = C.c_int32(0)
        check(_lib.lib().egsgpu_kernel_time(self.h, kernel, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def graph_stats(self):
        """The last graph-driven run: (graph launches, iterations with work,
        iterations the next run's graph holds) (egsgpu_graph_stats)."""
        v = [C.c_int32(0) for _ in range(3)]
        check(_lib.lib().egsgpu_graph_stats(self.h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def close(self):
        if self.h.value:
            _lib.lib().egsgpu_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_batched(sims, nhist):
    """egsgpu_run_batched: nhist[j] histories of each Simulation sims[j] in
    one event schedule (streams and slots of sims[0])."""
    hs = (C.c_void_p * len(sims))(*[s.h.value for s in sims])
    n = np.ascontiguousarray(nhist, dtype=np.int64)
    check(_lib.lib().egsgpu_run_batched(hs, len(sims), ptr(n)))


def dose_stats(sum_, sum2, n):
    """Per-history mean and its standard error (EGS_ScoringSingle::
    currentResult, egs_scoring.h:115-125)."""
    r = sum_ / n
    var = sum2 / n - r * r
    var = np.where(var > 0, var / (n - 1), 0.0)
    return r, np.sqrt(var)


def write_3ddose(path, xb, yb, zb, edep, dedep, mass, norm=1.602e-10):
    """.3ddose writer (EGS_DoseScoring::outputDoseFile,
    egs_dose_scoring.cpp:416-479): dims, bounds, dose = r*norm/mass,
    relative uncertainty (1 where r = 0); x fastest."""
    dose = edep * norm / mass
    rel = np.where(edep > 0, dedep / np.where(edep > 0, edep, 1), 1.0)
    with open(path, "w") as f:
        f.write("%d %d %d\n" % (len(xb) - 1, len(yb) - 1, len(zb) - 1))
        for b in (xb, yb, zb):
            f.write(" ".join("%g" % np.float32(v) for v in b) + " \n")
        f.write(" ".join("%g" % np.float32(v) for v in dose) + " \n")
        f.write(" ".join("%g" % np.float32(v) for v in rel) + " \n")

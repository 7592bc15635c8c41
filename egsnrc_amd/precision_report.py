"""Per-voxel report of what FP32 transport changes against FP64 transport.

The same workload is run twice on the device with the history schedule,
the same seeds and the same streams: once with the FP32 transport build and
once with the FP64 build (tallies are FP64 in both).  Both runs draw the
same random numbers, so their histories agree until rounding sends a
decision the other way; the per-voxel differences therefore measure the
effect of FP32 arithmetic on the dose, in units of the statistical
uncertainty.  The reference computes everything in FP64
(src/egsnrc.mortran, egs++/egs_scoring.h:63-125 for the sums).
"""
import numpy as np

from . import _lib
from .engine import Simulation, dose_stats


def run_pair(tables, xb, yb, zb, medium, rhor, source, transport, nhist,
             nstreams=16384, seeds=(1802, 9373), nbatch=10):
    """Run the workload in FP32 and FP64, each in nbatch launches of the
    same sizes (so that batch k of both runs holds the same histories);
    returns ((r32, s32), (r64, s64), n, t32, t64) with per-history mean
    deposits r, their standard errors s, and the per-batch total deposits
    t (MeV per history)."""
    out, tot = [], []
    sizes = [nhist // nbatch + (1 if k < nhist % nbatch else 0) for k in range(nbatch)]
    for fp64 in (0, 1):
        tp = _lib.EgsTransport.from_buffer_copy(transport)
        tp.fp64 = fp64
        sim = Simulation(tables, xb, yb, zb, medium, rhor, source, tp, nstreams=nstreams,
                         schedule="history")
        sim.seed(*seeds)
        t, prev = [], 0.0
        for m in sizes:
            sim.run(m)
            cur = float(sim.tallies()[0].sum())
            t.append((cur - prev) / m)
            prev = cur
        s, s2, n = sim.tallies()
        out.append(dose_stats(s, s2, n))
        tot.append(np.array(t))
        sim.close()
    return out[0], out[1], nhist, tot[0], tot[1]


def report(r32, s32, r64, s64, frac=0.1):
    """Summary over voxels above frac * max of the FP64 result:
    z = (r32 - r64) / sqrt(s32^2 + s64^2) per voxel."""
    m = r64 > frac * r64.max()
    den = np.sqrt(s32[m] ** 2 + s64[m] ** 2)
    ok = den > 0
    z = (r32[m][ok] - r64[m][ok]) / den[ok]
    rel = (r32[m] - r64[m]) / r64[m]
    return {
        "voxels": int(m.sum()),
        "chi2_per_dof": float(np.mean(z * z)) if z.size else 0.0,
        "frac_beyond_2sigma": float(np.mean(np.abs(z) > 2)) if z.size else 0.0,
        "max_abs_z": float(np.max(np.abs(z))) if z.size else 0.0,
        "mean_rel_diff": float(np.mean(rel)),
        "max_abs_rel_diff": float(np.max(np.abs(rel))),
        "mean_rel_sigma": float(np.mean(s64[m] / r64[m])),
        "total_rel_diff": float((r32.sum() - r64.sum()) / r64.sum()),
    }


def precision_report(tables, xb, yb, zb, medium, rhor, source, transport, nhist,
                     nstreams=16384, frac=0.1, nbatch=10):
    """The per-voxel report, plus the relative difference of the total
    deposit and its standard error from the paired per-batch differences
    (which carries both the correlation of a history's deposits over
    voxels and the correlation of the two runs)."""
    (r32, s32), (r64, s64), n, t32, t64 = run_pair(tables, xb, yb, zb, medium, rhor, source,
                                                    transport, nhist, nstreams, nbatch=nbatch)
    rep = report(r32, s32, r64, s64, frac)
    rep["histories"] = n
    d = (t32 - t64) / t64.mean()
    rep["total_rel_diff_sigma"] = float(d.std(ddof=1) / np.sqrt(len(d)))
    return rep, (r32, s32, r64, s64)

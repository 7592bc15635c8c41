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
             nstreams=16384, seeds=(1802, 9373)):
    """Run the workload in FP32 and FP64; returns ((r32, s32), (r64, s64),
    n) with per-history mean deposits r and their standard errors s."""
    out = []
    for fp64 in (0, 1):
        tp = _lib.EgsTransport.from_buffer_copy(transport)
        tp.fp64 = fp64
        sim = Simulation(tables, xb, yb, zb, medium, rhor, source, tp, nstreams=nstreams,
                         schedule="history")
        sim.seed(*seeds)
        sim.run(nhist)
        s, s2, n = sim.tallies()
        out.append(dose_stats(s, s2, n))
        sim.close()
    return out[0], out[1], nhist


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
                     nstreams=16384, frac=0.1):
    (r32, s32), (r64, s64), n = run_pair(tables, xb, yb, zb, medium, rhor, source,
                                          transport, nhist, nstreams)
    rep = report(r32, s32, r64, s64, frac)
    rep["histories"] = n
    return rep, (r32, s32, r64, s64)

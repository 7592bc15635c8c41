"""Statistical dose parity of two independent runs at the bar the engine
is held to: over the voxels above `frac` of the maximum dose, the
per-voxel z-scores z = (D_a - D_b) / sqrt(s_a^2 + s_b^2) must give

  * chi^2/dof <= 1.1, and
  * a fraction of |z| > 2 within 3 binomial standard deviations of the
    normal 4.55 %,

on at least `min_voxels` voxels (1800 by default: below that the 3-sigma
spread of chi^2/dof of correct runs, 3 sqrt(2/n), exceeds the 0.1 margin,
and the test has no power).  A lower `frac` than the 10 % of Dmax the
agreement is specified on only adds voxels, so it is a stricter check.

The normal 4.55 % holds when every voxel's mean and variance come from
many histories.  In small voxels at test sizes most voxels get their dose
from a few tens of histories, z then has heavier tails than a normal
variable (a Student t of few degrees of freedom) and two runs of the same
code exceed 2 sigma more often (5.1 % for 128^3 x 0.25 cm at 2e6 photon
histories, 4.8 % at 8e6).  Such a test passes `null`: a third run of the
same size, independent of both, and the fraction is then held to that of
the two reference runs (null vs a), within 3 standard deviations of the
difference of two binomial proportions.

Tests of the input front ends (a reference input file with its own small
phantom) judge fewer voxels; for them the chi^2 bound is the 0.1 % quantile
of chi^2 with nv degrees of freedom over nv whenever that is above 1.1
(it is 1.098 at 2000 voxels), so that every test has the same 1e-3 chance
of failing correct code."""
import numpy as np
from scipy import stats

from egsnrc_amd import workloads

CHI2_MAX = 1.1
P2 = 0.0455   # P(|z| > 2) for a standard normal


def chi2_bound(nv):
    return max(CHI2_MAX, float(stats.chi2.ppf(0.999, max(nv, 1))) / max(nv, 1))


def parity_stats(a, b, frac=0.1):
    return workloads.compare_doses(a, b, frac)


def z_scores(ma, sa, mb, sb, frac=0.1):
    """Per-voxel z of two results given as means and standard errors, over
    the voxels above frac of the larger maximum with a nonzero sigma."""
    keep = (np.maximum(ma, mb) > frac * max(ma.max(), mb.max())) & (sa ** 2 + sb ** 2 > 0)
    return (ma[keep] - mb[keep]) / np.sqrt(sa[keep] ** 2 + sb[keep] ** 2)


def check_parity_means(ma, sa, mb, sb, label="", min_voxels=1800, frac=0.1):
    """The bar of check_parity for results given as means and standard
    errors (e.g. read back from .3ddose files)."""
    z = z_scores(ma, sa, mb, sb, frac)
    nv = z.size
    chi2, out2 = float(np.mean(z * z)), float(np.mean(np.abs(z) > 2))
    band = 3 * np.sqrt(P2 * (1 - P2) / max(nv, 1))
    msg = ("%s: chi2/dof %.4f  >2sigma %.4f (allowed %.4f..%.4f)  voxels %d"
           % (label, chi2, out2, P2 - band, P2 + band, nv))
    print(msg)
    assert nv >= min_voxels, msg
    assert chi2 <= chi2_bound(nv), msg
    assert abs(out2 - P2) <= band, msg
    return chi2, out2, nv


def check_parity(a, b, counters=None, min_voxels=1800, label="", frac=0.1,
                 total_rtol=None, null=None):
    """a, b = (sum, sum2, nhist).  Asserts the bar above, no stack, bank or
    tally-table overflow in `counters` (engine counters 5 and 6) and, with
    total_rtol, agreement of the total deposit per history."""
    chi2, out2, nv = workloads.compare_doses(a, b, frac)
    p2 = P2
    band = 3 * np.sqrt(P2 * (1 - P2) / max(nv, 1))
    if null is not None:
        chi2n, p2, nvn = workloads.compare_doses(a, null, frac)
        band = 3 * np.sqrt(p2 * (1 - p2) * (1.0 / max(nv, 1) + 1.0 / max(nvn, 1)))
        print("%s null (reference vs reference): chi2/dof %.4f  >2sigma %.4f  voxels %d"
              % (label, chi2n, p2, nvn))
    msg = ("%s: chi2/dof %.4f  >2sigma %.4f (allowed %.4f..%.4f)  voxels %d"
           % (label, chi2, out2, p2 - band, p2 + band, nv))
    if counters is not None:
        msg += "  overflows %d/%d" % (counters[5], counters[6])
    print(msg)
    if counters is not None:
        assert counters[5] == 0, "stack or bank overflow"
        assert counters[6] == 0, "history tally table overflow"
    assert nv >= min_voxels, msg
    assert chi2 <= chi2_bound(nv), msg
    assert abs(out2 - p2) <= band, msg
    if total_rtol is not None:
        ta, tb = a[0].sum() / a[2], b[0].sum() / b[2]
        assert abs(ta - tb) <= total_rtol * ta, (ta, tb)
    return chi2, out2, nv

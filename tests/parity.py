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
difference of two binomial proportions.  The same tails make chi^2/dof of
correct runs both larger than 1 and far more variable than 2/nv suggests
(two runs of C0 at 6e6 histories gave 1.004 and 1.107 over 2700 voxels,
while 1e7-history depth doses agree with the oracle to 0.03 %); with
`null`, chi^2/dof may exceed 1.1 only as far as the two reference runs'
does, within 3 standard errors of the difference, the standard errors
taken from the spread of z^2 itself.

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


def voxel_z(a, b, frac=0.1):
    """Per-voxel z of two runs a, b = (sum, sum2, nhist) over the voxels
    above frac of the larger maximum with a nonzero variance."""
    def st(t):
        s, s2, n = t
        r = s / n
        v = s2 / n - r * r
        return r, np.where(v > 0, v / (n - 1), 0.0)
    ra, va = st(a)
    rb, vb = st(b)
    m = (np.maximum(ra, rb) > frac * max(ra.max(), rb.max())) & (va + vb > 0)
    return (ra[m] - rb[m]) / np.sqrt(va[m] + vb[m])


def z_scores(ma, sa, mb, sb, frac=0.1):
    """Per-voxel z of two results given as means and standard errors, over
    the voxels above frac of the larger maximum with a nonzero sigma."""
    keep = (np.maximum(ma, mb) > frac * max(ma.max(), mb.max())) & (sa ** 2 + sb ** 2 > 0)
    return (ma[keep] - mb[keep]) / np.sqrt(sa[keep] ** 2 + sb[keep] ** 2)


def _judge(z, label, min_voxels, zn=None, extra=""):
    nv = z.size
    chi2, out2 = float(np.mean(z * z)), float(np.mean(np.abs(z) > 2))
    chi2_max, p2 = chi2_bound(nv), P2
    band = 3 * np.sqrt(P2 * (1 - P2) / max(nv, 1))
    if zn is not None:
        # calibrated on two runs of the same reference: chi^2/dof may exceed
        # 1.1 only as far as theirs does, within 3 standard errors of the
        # difference (from the spread of z^2, which carries the heavy tails)
        nvn = zn.size
        chi2n, p2 = float(np.mean(zn * zn)), float(np.mean(np.abs(zn) > 2))
        se = np.sqrt(np.var(z * z) / max(nv, 1) + np.var(zn * zn) / max(nvn, 1))
        chi2_max = max(chi2_max, chi2n + 3 * se)
        band = 3 * np.sqrt(p2 * (1 - p2) * (1.0 / max(nv, 1) + 1.0 / max(nvn, 1)))
        print("%s null (reference vs reference): chi2/dof %.4f  >2sigma %.4f  voxels %d"
              % (label, chi2n, p2, nvn))
    msg = ("%s: chi2/dof %.4f (max %.4f)  >2sigma %.4f (allowed %.4f..%.4f)  voxels %d%s"
           % (label, chi2, chi2_max, out2, p2 - band, p2 + band, nv, extra))
    print(msg)
    assert nv >= min_voxels, msg
    assert chi2 <= chi2_max, msg
    assert abs(out2 - p2) <= band, msg
    return chi2, out2, nv


def check_parity_means(ma, sa, mb, sb, label="", min_voxels=1800, frac=0.1, null=None):
    """The bar of check_parity for results given as means and standard
    errors (e.g. read back from .3ddose files); null = (mean, sigma) of a
    second reference run."""
    zn = z_scores(ma, sa, null[0], null[1], frac) if null is not None else None
    return _judge(z_scores(ma, sa, mb, sb, frac), label, min_voxels, zn)


def check_parity(a, b, counters=None, min_voxels=1800, label="", frac=0.1,
                 total_rtol=None, null=None):
    """a, b = (sum, sum2, nhist): a the reference run, b the run judged.
    Asserts the bar above, no stack, bank or tally-table overflow in
    `counters` (engine counters 5 and 6) and, with total_rtol, agreement
    of the total deposit per history.  null: a second reference run,
    independent of a and b, of the same size as a."""
    extra = ""
    if counters is not None:
        extra = "  overflows %d/%d" % (counters[5], counters[6])
        assert counters[5] == 0, "stack or bank overflow"
        assert counters[6] == 0, "history tally table overflow"
    zn = voxel_z(a, null, frac) if null is not None else None
    r = _judge(voxel_z(a, b, frac), label, min_voxels, zn, extra)
    if total_rtol is not None:
        ta, tb = a[0].sum() / a[2], b[0].sum() / b[2]
        assert abs(ta - tb) <= total_rtol * ta, (ta, tb)
    return r

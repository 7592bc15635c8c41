"""Synthetic phantoms and sources for the configurations the engine is
measured on (shared by tests and bench.py)."""
import numpy as np
from scipy.ndimage import gaussian_filter


def grf_density(n, sigma, seed, lo=0.2, hi=2.0):
    """Smoothed Gaussian random field mapped linearly to [lo, hi] g/cm^3."""
    rng = np.random.default_rng(seed)
    f = gaussian_filter(rng.normal(size=(n, n, n)), sigma, mode="wrap")
    f = (f - f.mean()) / f.std()
    rho = 1.1 + 0.45 * f
    return np.clip(rho, lo, hi)


def kerma_phantom(n=64, dx=0.4, seed=4):
    """Photon-only kerma phantom: n^3 voxels of water with a smoothed random
    density field clipped to [0.2, 2.0] (voxel index z-major, x fastest)."""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    rho = grf_density(n, 3.0, seed)           # array [ix, iy, iz]
    rhor = np.transpose(rho, (2, 1, 0)).reshape(-1)   # x fastest
    med = np.zeros(n ** 3, dtype=np.int32)
    return xb, xb.copy(), zb, med, rhor


def slab_phantom(n=128, dx=0.25):
    """Heterogeneous slab phantom: water / lung / bone / water along z
    (media 0 = water, 1 = lung, 2 = bone of the 521icru table blob)."""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    zc = 0.5 * (zb[1:] + zb[:-1])
    mz = np.where(zc < 8, 0, np.where(zc < 16, 1, np.where(zc < 20, 2, 0)))
    med = np.broadcast_to(mz[:, None, None], (n, n, n)).reshape(-1).astype(np.int32)
    return xb, xb.copy(), zb, med, None


def falling_spectrum(emin=0.1, emax=6.0, nbins=60):
    """60-bin histogram with fluence falling linearly from emin to emax."""
    e = np.linspace(emin, emax, nbins + 1)
    c = 0.5 * (e[1:] + e[:-1])
    return e, (emax - c)


def compare_doses(a, b, frac=0.1):
    """Voxel-wise statistical agreement of two independent runs.
    a, b = (sum, sum2, n).  Returns (chi2/dof, fraction of voxels outside
    2 sigma, number of voxels) over voxels above frac * Dmax."""
    def stats(t):
        s, s2, n = t
        r = s / n
        v = s2 / n - r * r
        v = np.where(v > 0, v / (n - 1), 0.0)
        return r, v
    ra, va = stats(a)
    rb, vb = stats(b)
    dmax = max(ra.max(), rb.max())
    m = (np.maximum(ra, rb) > frac * dmax) & (va + vb > 0)
    z = (ra[m] - rb[m]) / np.sqrt(va[m] + vb[m])
    return float(np.mean(z * z)), float(np.mean(np.abs(z) > 2)), int(m.sum())

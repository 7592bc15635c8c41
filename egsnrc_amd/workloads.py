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


def water_cube(n=30, dx=1.0):
    """C0 phantom: n^3 voxels of water, x/y centred, z from 0 (medium 0)."""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    return xb, xb.copy(), zb, np.zeros(n ** 3, dtype=np.int32), None


def al_cube_phantom(n=256, dx=0.2, side=3.0, depth=3.0, al=4):
    """C2 phantom: n^3 voxels of water with an aluminium cube of
    round(side/dx) voxels per edge (15 at 2 mm), laterally the voxels
    n/2 - k/2 - 1 ... (centred to within half a voxel), from z = depth
    (media of the 521icru blob: 0 water, 4 aluminium)."""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    k = int(round(side / dx))
    lo = n // 2 - (k + 1) // 2
    inx = np.zeros(n, bool)
    inx[lo:lo + k] = True
    iz0 = int(round(depth / dx))
    inz = np.zeros(n, bool)
    inz[iz0:iz0 + k] = True
    m = (inz[:, None, None] & inx[None, :, None] & inx[None, None, :]).astype(np.int32) * al
    return xb, xb.copy(), zb, m.reshape(-1), None


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


def batch_jobs(njobs=256, seed=3, hist_scale=1.0, nmin=32, nmax=160):
    """The many-small-jobs workload: njobs independent simulations drawn
    from numpy default_rng(seed).  Per job: cube edge n ~ U{nmin..nmax}
    voxels of 1-5 mm; a Gaussian random field (sigma 3 voxels) thresholded
    at fixed quantiles into air / lung / water / bone (5 / 20 / 60 / 15 %)
    with +-5 % uniform density noise; even jobs a photon beam with the
    falling 0.1-6 MeV spectrum, odd jobs a 4-20 MeV electron beam; the beam
    covers the central half of the entrance face; histories log-uniform in
    [1e5, 1e7] times hist_scale.  Media indices are those of the 521icru
    blob (0 water, 1 lung, 2 bone, 3 air)."""
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(njobs):
        n = int(rng.integers(nmin, nmax + 1))
        dx = float(rng.uniform(0.1, 0.5))
        f = gaussian_filter(rng.normal(size=(n, n, n)).astype(np.float32), 3.0, mode="wrap")
        q = np.quantile(f, [0.05, 0.25, 0.85])
        m = np.select([f < q[0], f < q[1], f < q[2]], [3, 1, 0], 2).astype(np.int32)
        rho = 1.0 + rng.uniform(-0.05, 0.05, size=f.shape)
        med = np.ascontiguousarray(np.transpose(m, (2, 1, 0))).reshape(-1)
        rhor = np.ascontiguousarray(np.transpose(rho, (2, 1, 0))).reshape(-1)
        half = n * dx / 2
        xb = np.linspace(-half, half, n + 1)
        zb = np.linspace(0.0, n * dx, n + 1)
        if j % 2 == 0:
            src = dict(charge=0, energy=0.0, spectrum=falling_spectrum())
        else:
            src = dict(charge=-1, energy=float(rng.uniform(4.0, 20.0)), spectrum=None)
        src.update(x0=-half / 2, x1=half / 2, y0=-half / 2, y1=half / 2, z=-1.0)
        nh = int(round(10 ** rng.uniform(5.0, 7.0) * hist_scale))
        jobs.append(dict(xb=xb, yb=xb.copy(), zb=zb, med=med, rhor=rhor, src=src,
                         nhist=max(1, nh), n=n, dx=dx))
    return jobs

"""Fano test of the charged-particle transport (egs_fano_source,
egs++/sources/egs_fano_source/egs_fano_source.h): in a medium of one
composition whose density varies from voxel to voxel, particles that start
uniformly per unit mass with isotropic directions give a uniform dose where
charged-particle equilibrium holds, equal to the emitted energy per unit
mass, whatever the density pattern (Fano's theorem).  Per-mass interaction
data are the same at every density (the engine scales one medium's tables
by the voxel's relative density), and photons are deposited where they are
made (PCUT above the source energy), so the theorem applies exactly; what it
checks end to end is the density scaling of every step length and of the
multiple scattering, and the exact boundary crossing at every voxel face
where the density jumps."""
import os

import numpy as np
import pytest

from egsnrc_amd import _lib
from egsnrc_amd.engine import Simulation, Tables, make_fano_source
from egsnrc_amd import workloads
from tests.parity import chi2_bound

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BLOB = os.path.join(ROOT, "egsnrc_amd", "data", "521icru.egstab")


def fano_case(n=24, dx=0.25, energy=1.0, lo=0.5, hi=2.0, seed=11):
    """n^3 voxels of water with relative densities from a smoothed random
    field in [lo, hi]; 1 MeV electrons from a Fano source filling the
    phantom; ECUT 0.521 MeV, PCUT 2 MeV (above every photon energy)."""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    rho = workloads.grf_density(n, 2.0, seed, lo=lo, hi=hi)
    rhor = np.ascontiguousarray(np.transpose(rho, (2, 1, 0))).reshape(-1)
    med = np.zeros(n ** 3, dtype=np.int32)
    box = (xb[0], xb[-1], xb[0], xb[-1], zb[0], zb[-1])
    src = make_fano_source(-1, energy, box, rho_max=float(rhor.max()))
    tp = _lib.default_transport(ecut=0.521, pcut=2.0, kerma=0)
    return xb, xb.copy(), zb, med, rhor, src, tp


@pytest.mark.gpu
def test_fano_source_rejects_a_low_rho_max(engine):
    """The acceptance ratio rho/rho_max may not exceed 1 in any voxel."""
    xb, yb, zb, med, rhor, src, tp = fano_case(n=8)
    src[0].rho_max = 0.9 * float(rhor.max())
    with pytest.raises(_lib.EngineError, match="rho_max"):
        Simulation(Tables(BLOB), xb, yb, zb, med, rhor, src, tp)


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", ["history", "event"])
def test_fano_dose_is_uniform(engine, schedule):
    """Over the voxels farther than the largest electron range from the
    faces (CSDA range of 1 MeV electrons in water 0.44 g/cm^2, 0.9 cm at
    the lowest density): the per-voxel z of D_v against E0/M gives
    chi^2/dof within the 0.1 % quantile of chi^2, the mean dose agrees with
    E0/M within 3 standard errors, and z is uncorrelated with the voxel
    density (|r| < 3/sqrt(n))."""
    n, dx = 24, 0.25
    xb, yb, zb, med, rhor, src, tp = fano_case(n, dx)
    N = 20_000_000
    sim = Simulation(Tables(BLOB), xb, yb, zb, med, rhor, src, tp, schedule=schedule)
    sim.run(N)
    s, s2, nh = sim.tallies()
    c = sim.counters()
    assert nh == N and c[5] == 0 and c[6] == 0
    vol = dx ** 3
    mass = vol * rhor * 1.0          # water of the 521icru blob: 1 g/cm^3
    e0 = 1.0
    dexp = e0 / mass.sum()           # MeV/g per history
    r = s / N
    v = np.maximum(s2 / N - r * r, 0) / (N - 1)
    d, sd = r / mass, np.sqrt(v) / mass
    k = 5                            # 1.25 cm > 0.9 cm from every face
    idx = np.arange(n)
    inner = (idx >= k) & (idx < n - k)
    m = (inner[None, None, :] & inner[None, :, None] & inner[:, None, None]).reshape(-1)
    z = (d[m] - dexp) / sd[m]
    nv = z.size
    chi2 = float(np.mean(z * z))
    mean_rel = float(np.mean(d[m]) / dexp - 1)
    sem_rel = float(np.sqrt(np.sum(sd[m] ** 2)) / nv / dexp)
    corr = float(np.corrcoef(z, rhor[m])[0, 1])
    print("Fano %s: %d voxels, chi2/dof %.4f (max %.4f), mean dose / (E0/M) - 1 = %.2e "
          "(sem %.1e), corr(z, rho) %.4f" % (schedule, nv, chi2, chi2_bound(nv), mean_rel,
                                            sem_rel, corr))
    assert chi2 <= chi2_bound(nv)
    assert abs(mean_rel) < 3 * sem_rel
    assert abs(corr) < 3 / np.sqrt(nv)

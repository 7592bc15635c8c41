"""Small end-to-end invocation of the engine, used by the smoke entry point:
device voxel tracing and one small kerma transport run, both checked
against the CPU oracle."""
import os


def run():
    import numpy as np
    from egsnrc_amd import _lib
    from egsnrc_amd.engine import Simulation, Tables, make_source
    from egsnrc_amd import workloads
    from tests import oracle_lib
    xb = np.linspace(-2, 2, 9)
    g, keep = _lib.make_geometry(xb, xb, np.linspace(0, 4, 9))
    x = np.array([[0.1, 0.2, -1.0], [-3.0, 0.3, 1.1]])
    u = np.array([[0.1, 0.2, 0.97], [1.0, 0.0, 0.0]])
    u /= np.linalg.norm(u, axis=1)[:, None]
    ra, sa, na = _lib.trace_rays(g, x, u, 64, 32)
    rb, sb, nb = oracle_lib.trace_rays(g, x, u, 64, 32)
    assert np.array_equal(na, nb), "voxel trace length mismatch"
    for r in range(len(x)):
        k = na[r]
        assert np.array_equal(ra[r, :k], rb[r, :k]) and np.array_equal(sa[r, :k], sb[r, :k]), \
            "voxel trace mismatch"
    # tiny transport: 8^3 kerma phantom, 2 MeV photons
    xb, yb, zb, med, rhor = workloads.kerma_phantom(8, 2.0)
    src = make_source(0, 2.0, x0=-4, x1=4, y0=-4, y1=4, z=-1.0)
    tp = _lib.default_transport(ecut=20.0, kerma=1)
    blob = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "egsnrc_amd",
                        "data", "521icru.egstab")
    sim = Simulation(Tables(blob), xb, yb, zb, med, rhor, src, tp, nstreams=4096)
    sim.run(100000)
    s, s2, n = sim.tallies()
    g2, keep2 = _lib.make_geometry(xb, yb, zb, med, rhor)
    o = oracle_lib.OracleSimulation(blob, g2, src[0], tp, nstreams=64, nthreads=4)
    o.run(100000)
    so, s2o, no = o.tallies()
    ta, tb = s.sum() / n, so.sum() / no
    assert abs(ta - tb) < 0.02 * tb, (ta, tb)
    chi2, out2, nv = workloads.compare_doses((s, s2, n), (so, s2o, no))
    assert chi2 < 1.5, chi2
    print("smoke ok: edep/hist gpu %.5f oracle %.5f chi2/dof %.3f (%d voxels)" % (ta, tb, chi2, nv))

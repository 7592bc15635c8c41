"""The C++ host application (egs++-style .egsinp front end over the C ABI)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "egsnrc_amd", "egs_gpu_app")
EX = os.path.join(ROOT, "egsnrc_amd", "examples", "c0_pencil_water.egsinp")
BLOB700 = os.path.join(ROOT, "egsnrc_amd", "data", "700icru.egstab")


def read_3ddose(path):
    tok = open(path).read().split()
    nx, ny, nz = int(tok[0]), int(tok[1]), int(tok[2])
    k = 3
    b = []
    for n in (nx, ny, nz):
        b.append(np.array(tok[k:k + n + 1], float)); k += n + 1
    nv = nx * ny * nz
    d = np.array(tok[k:k + nv], float); k += nv
    e = np.array(tok[k:k + nv], float)
    return (nx, ny, nz), b, d, e


def test_app_rejects_bad_input(tmp_path):
    bad = tmp_path / "bad.egsinp"
    bad.write_text(":start geometry definition:\n x = 1\n")
    r = subprocess.run([APP, "-i", str(bad), "-p", BLOB700], capture_output=True, text=True)
    assert r.returncode != 0
    assert "stop" in r.stderr.lower()


def test_app_rejects_unsupported_geometry(tmp_path):
    txt = open(EX).read().replace("EGS_XYZGeometry", "EGS_cSpheres")
    f = tmp_path / "sph.egsinp"
    f.write_text(txt)
    r = subprocess.run([APP, "-i", str(f), "-p", BLOB700], capture_output=True, text=True)
    assert r.returncode != 0 and "EGS_XYZGeometry" in r.stderr


@pytest.mark.gpu
def test_app_runs_c0_input(tmp_path):
    txt = open(EX).read().replace("ncase  = 1000000", "ncase  = 200000")
    f = tmp_path / "c0.egsinp"
    f.write_text(txt)
    r = subprocess.run([APP, "-i", str(f), "-p", BLOB700], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    dims, b, d, e = read_3ddose(str(tmp_path / "c0.3ddose"))
    assert dims == (30, 30, 30)
    assert np.allclose(b[0], np.arange(-15, 16))
    d3 = d.reshape(30, 30, 30)       # [iz, iy, ix]
    # the pencil beam enters at x = y = 0, a voxel boundary: isWhere puts a
    # point on a plane in the lower voxel (findRegion), i.e. ix = iy = 14
    axis = d3[:, 14, 14]
    assert axis.max() == d.max()
    assert np.all(axis[:25] > 0)
    # dose = r * 1.602e-10 / (V rho) in Gy per history with the relative
    # uncertainty beside it (egs_dose_scoring.cpp:416-479): the same problem
    # run through the C ABI from Python agrees voxel by voxel, and in the
    # energy deposited per history
    from egsnrc_amd import _lib, workloads
    from egsnrc_amd.engine import Simulation, Tables, make_source
    from tests.parity import check_parity_means
    xb, yb, zb, med, rhor = workloads.water_cube(30, 1.0)
    src = make_source(0, 1.25, shape="point", x0=0.0, y0=0.0, z=-1.0)
    tp = _lib.default_transport(ecut=0.7, pcut=0.01, rayleigh=0)
    mass = 1.0 * 1.0                      # 1 cm^3 of water at 1 g/cm^3

    def python_run(seeds):
        sim = Simulation(Tables(BLOB700), xb, yb, zb, med, rhor, src, tp)
        sim.seed(*seeds)
        sim.run(200000)
        su, s2, n = sim.tallies()
        sim.close()
        r = su / n
        var = np.maximum(s2 / n - r * r, 0) / (n - 1)
        return r * 1.602e-10 / mass, np.sqrt(var) * 1.602e-10 / mass
    dg, sg = python_run((4711, 20000))
    null = python_run((9001, 30000))      # the spread of two runs of the same problem
    sa = d * e                            # the file's relative uncertainty
    check_parity_means(dg, sg, d, sa, label="c0 .3ddose vs C ABI", frac=0.002,
                       min_voxels=400, null=null)
    ea, eg = d.sum() / 1.602e-10, dg.sum() / 1.602e-10     # MeV per history
    assert abs(ea - eg) < 0.01 * eg, (ea, eg)   # ~0.1 % statistical; a wrong norm is far off

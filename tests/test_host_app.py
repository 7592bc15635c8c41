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

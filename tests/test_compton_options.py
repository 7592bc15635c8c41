"""The bound Compton options (IBCMP; egs_advanced_application.cpp:497-501:
Off 0, On 1, Simple 2, norej 3) and the other on/off photon options as the
host app reads them; GPU dose parity of the engine with the oracle for the
two options the other tests leave out, On (bound Compton with the
impulse-approximation rejection) and Simple (egsnrc.mortran:528-806)."""
import os

import numpy as np
import pytest

from egsnrc_amd import _lib
from egsnrc_amd.engine import Simulation, Tables, make_source
from tests import oracle_lib
from tests.parity import check_parity

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BLOB = os.path.join(ROOT, "egsnrc_amd", "data", "521icru.egstab")


def test_host_app_maps_the_photon_options(tmp_path):
    from tests.test_dosxyz_frontend import EXAMPLE, parse, transport_options
    txt = open(EXAMPLE).read()
    for name, k in (("Off", 0), ("On", 1), ("Simple", 2), ("norej", 3)):
        f = tmp_path / ("bc_%s.egsinp" % name)
        f.write_text(txt.replace("Bound Compton scattering= Off",
                                 "Bound Compton scattering= " + name))
        p, _ = parse(f)
        assert transport_options(p)["bound_compton"] == str(k)
    f = tmp_path / "relax.egsinp"
    f.write_text(txt.replace("Atomic relaxations= Off", "Atomic relaxations= eadl"))
    p, _ = parse(f)
    assert transport_options(p)["relax"] == "1"
    # the pre-EADL relaxation and custom Rayleigh form factors are refused
    import subprocess
    from tests.test_host_app import APP
    for old, new in (("Atomic relaxations= Off", "Atomic relaxations= simple"),
                     ("Rayleigh scattering= Off", "Rayleigh scattering= custom")):
        f.write_text(txt.replace(old, new))
        r = subprocess.run([APP, "-i", str(f), "--parse-only"], capture_output=True, text=True,
                           timeout=60)
        assert r.returncode != 0 and "not supported" in r.stderr, r.stderr


def compton_case(ibcmp, n=24, dx=0.5, energy=0.2, half=3.0):
    """200 keV photons (6 x 6 cm^2) on water with a bone slab at 2-4 cm
    depth in 5 mm voxels, bound Compton option ibcmp."""
    xb = np.linspace(-n * dx / 2, n * dx / 2, n + 1)
    zb = np.linspace(0.0, n * dx, n + 1)
    zc = 0.5 * (zb[1:] + zb[:-1])
    mz = np.where((zc > 2.0) & (zc < 4.0), 2, 0)
    med = np.broadcast_to(mz[:, None, None], (n, n, n)).reshape(-1).astype(np.int32)
    src = make_source(0, energy, x0=-half, x1=half, y0=-half, y1=half, z=-1.0)
    tp = _lib.default_transport(kerma=0, ecut=0.521, pcut=0.001, bound_compton=ibcmp)
    return xb, xb.copy(), zb, med, None, src, tp


HIST = 1000000
_ORACLE = {}


def _oracle(ibcmp, seeds=None):
    key = (ibcmp, seeds)
    if key not in _ORACLE:
        xb, yb, zb, med, rhor, src, tp = compton_case(ibcmp)
        g, keep = _lib.make_geometry(xb, yb, zb, med, rhor)
        o = oracle_lib.OracleSimulation(BLOB, g, src[0], tp, nstreams=256,
                                        nthreads=max(1, min(os.cpu_count() or 1, 16)))
        if seeds:
            o.seed(*seeds)
        o.run(HIST)
        _ORACLE[key] = o.tallies()
    return _ORACLE[key]


@pytest.mark.gpu
@pytest.mark.parametrize("ibcmp", [1, 2])
def test_bound_compton_option_dose_parity(engine, ibcmp):
    xb, yb, zb, med, rhor, src, tp = compton_case(ibcmp)
    sim = Simulation(Tables(BLOB), xb, yb, zb, med, rhor, src, tp, nstreams=8192)
    sim.run(HIST)
    b = sim.tallies()
    assert b[2] == HIST
    c = sim.counters()
    assert c[5] == 0 and c[6] == 0
    check_parity(_oracle(ibcmp), b, c, label="bound compton %d" % ibcmp,
                 null=_oracle(ibcmp, (4242, 2718)))

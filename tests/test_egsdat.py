"""egs++ .egsdat state files in the host application: the layout that
EGS_Application::outputData / readData / addState write and read
(egs++/egs_application.cpp:482-589; run control egs_run_control.cpp:149-175,
ranmar egs_rndm.cpp:83-131 and 265-281, source egs_base_source.h:459-492,
dose scoring egs_dose_scoring.cpp:481-540, scoring arrays
egs_scoring.h:144-165 and 316-365, 64-bit counts egs_functions.cpp:58-80),
and the calculation modes of the run control (first, restart, analyze,
combine; egs_run_control.cpp:111-116, 185-215; combineResults
egs_application.cpp:624-675).

The CPU tests write state files the way the reference's writers lay them
out, with default 6-digit stream formatting, and check that combine and
analyze turn them into the .3ddose the sums imply; the GPU test runs,
restarts and checks the accumulated state."""
import os
import subprocess

import numpy as np
import pytest

from tests.test_host_app import APP, BLOB700, EX, read_3ddose

MEV_J = 1.602e-10


def i64(n):
    return " %d %d" % (n // 1000000000, n % 1000000000)


def reference_layout_egsdat(path, ndone, sums, sums2, narrays=2, cpu=12.5):
    """A state file laid out as an egs++ application with a parallel beam
    and an egs_dose_scoring object (region dose + output dose file)
    writes it."""
    out = [i64(ndone) + "  %g" % cpu, i64(ndone)]
    out.append(i64(12345) + " 128  128")
    out.append(" ".join("%g" % (k / 128.0) for k in range(128)) + " ")
    out.append("96 32 362436 1802 9373 0")
    out.append(" ".join(str(7 * k + 1) for k in range(97)) + " ")
    out.append(i64(ndone) + i64(ndone) + " %g %g" % (1.25 * ndone, 1.5625 * ndone))
    out.append(i64(ndone))
    for _ in range(narrays):
        out.append("%d  %d" % (len(sums), ndone % 65536))
        out.append(i64(ndone) + i64(ndone - ndone % 65536))
        for s, s2 in zip(sums, sums2):
            out.append("%d  %g  %g" % (ndone, s, s2))
    path.write_text("\n".join(out) + "\n")


def small_input(tmp_path, calculation):
    txt = open(EX).read()
    txt = txt.replace("ncase  = 1000000", "ncase  = 1000000\n    calculation = %s" % calculation)
    f = tmp_path / "job.egsinp"
    f.write_text(txt)
    return f


def sums_for(seed, nvox, n):
    rng = np.random.default_rng(seed)
    s = rng.uniform(0, 5, nvox) * n * 1e-3
    s2 = s * s / n * rng.uniform(1.5, 3, nvox)
    return s, s2


def test_combine_parallel_job_state_files(tmp_path):
    """calculation = combine sums <output>_w<j>.egsdat (jobs 1 and 3 here,
    a gap at 2) and writes the .3ddose of the summed tallies and the
    combined state file."""
    f = small_input(tmp_path, "combine")
    nvox = 30 ** 3
    (sa, s2a), (sb, s2b) = sums_for(1, nvox, 4000), sums_for(2, nvox, 6000)
    reference_layout_egsdat(tmp_path / "job_w1.egsdat", 4000, sa, s2a)
    reference_layout_egsdat(tmp_path / "job_w3.egsdat", 6000, sb, s2b)
    r = subprocess.run([APP, "-i", str(f), "-p", BLOB700], capture_output=True, text=True,
                       cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "job_w1.egsdat" in r.stdout and "job_w3.egsdat" in r.stdout
    dims, b, d, e = read_3ddose(str(tmp_path / "job.3ddose"))
    n = 10000
    # the file carries 6 significant digits, as the reference writes
    s = np.array(["%g" % v for v in sa], float) + np.array(["%g" % v for v in sb], float)
    s2 = np.array(["%g" % v for v in s2a], float) + np.array(["%g" % v for v in s2b], float)
    mass = 1.0   # 1 cm voxels of water (1 g/cm3)
    dose = (s / n) * MEV_J / mass
    r_ = s / n
    var = np.maximum(s2 / n - r_ * r_, 0) / (n - 1)
    assert np.allclose(d, dose, rtol=2e-6)
    assert np.allclose(e, np.sqrt(var) / r_, rtol=1e-5)
    # the combined state: 10000 histories, the sums of both files
    txt = (tmp_path / "job.egsdat").read_text().split()
    assert int(txt[0]) * 10 ** 9 + int(txt[1]) == n


def test_analyze_reads_the_state_file(tmp_path):
    f = small_input(tmp_path, "analyze")
    nvox = 30 ** 3
    s, s2 = sums_for(3, nvox, 5000)
    reference_layout_egsdat(tmp_path / "job.egsdat", 5000, s, s2)
    r = subprocess.run([APP, "-i", str(f), "-p", BLOB700], capture_output=True, text=True,
                       cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    dims, b, d, e = read_3ddose(str(tmp_path / "job.3ddose"))
    s6 = np.array(["%g" % v for v in s], float)
    assert np.allclose(d, s6 / 5000 * MEV_J, rtol=2e-6)


def test_truncated_state_file_is_an_error(tmp_path):
    f = small_input(tmp_path, "analyze")
    s, s2 = sums_for(4, 30 ** 3, 100)
    reference_layout_egsdat(tmp_path / "job.egsdat", 100, s, s2)
    txt = (tmp_path / "job.egsdat").read_text()
    (tmp_path / "job.egsdat").write_text(txt[: len(txt) // 2])
    r = subprocess.run([APP, "-i", str(f), "-p", BLOB700], capture_output=True, text=True,
                       cwd=tmp_path, timeout=120)
    assert r.returncode != 0 and "job.egsdat" in r.stderr


@pytest.mark.gpu
def test_run_then_restart_accumulates(tmp_path):
    """A first run writes the state files; calculation = restart continues
    from the engine checkpoint (every stream exactly), and, with the
    checkpoint removed, from the .egsdat sums alone: the history count
    adds up and the dose stays that of the first run within statistics."""
    txt = open(EX).read().replace("ncase  = 1000000", "ncase  = 200000")
    f = tmp_path / "c0.egsinp"
    f.write_text(txt)
    run = lambda: subprocess.run([APP, "-i", str(f), "-p", BLOB700], capture_output=True,
                                 text=True, cwd=tmp_path, timeout=300)
    r = run()
    assert r.returncode == 0, r.stderr
    d1 = read_3ddose(str(tmp_path / "c0.3ddose"))
    assert (tmp_path / "c0.egsdat").exists() and (tmp_path / "c0.egsgpu").exists()
    f.write_text(txt.replace("ncase  = 200000", "ncase  = 200000\n    calculation = restart"))
    r = run()
    assert r.returncode == 0, r.stderr
    assert "restart from" in r.stdout and ".egsgpu" in r.stdout
    tok = (tmp_path / "c0.egsdat").read_text().split()
    assert int(tok[0]) * 10 ** 9 + int(tok[1]) == 400000
    os.remove(tmp_path / "c0.egsgpu")
    r = run()
    assert r.returncode == 0, r.stderr
    assert "c0.egsdat" in r.stdout
    tok = (tmp_path / "c0.egsdat").read_text().split()
    assert int(tok[0]) * 10 ** 9 + int(tok[1]) == 600000
    d3 = read_3ddose(str(tmp_path / "c0.3ddose"))
    m = d1[2] > 0.2 * d1[2].max()
    z = (d1[2][m] - d3[2][m]) / np.sqrt((d1[2][m] * d1[3][m]) ** 2 + (d3[2][m] * d3[3][m]) ** 2)
    assert np.mean(z * z) < 1.5 and m.sum() > 50

"""Build the gfx950 engine library ``egsnrc_amd/libegsgpu.so`` with hipcc.

Every ``csrc/*.hip`` and ``csrc/*.cpp`` file is compiled for gfx950 into
``egsnrc_amd/_build`` and linked into one shared library.  Run as
``python -m egsnrc_amd.build`` or through ``__graft_entry__.build()``.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libegsgpu.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

# -packed-fp32-ops: keep the compiler from emitting packed FP32 VALU ops.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-I", os.path.join(HERE, "..", "include"),
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "hip"] + HIP_FLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


HOST_APP = os.path.join(HERE, "egs_gpu_app")


def build_app():
    """The C++ host application linked against libegsgpu.so."""
    srcs = [os.path.join(HERE, "host", f)
            for f in ("egs_input.cpp", "egs_phantom.cpp", "egs_phsp.cpp", "egs_dat.cpp",
                      "dosxyz_input.cpp", "egs_gpu_app.cpp")]
    cmd = ["g++", "-O2", "-std=c++17", "-o", HOST_APP + ".tmp"] + srcs + [
        "-I", os.path.join(HERE, "..", "include"), "-L", HERE, "-legsgpu", "-lz",
        "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host app build failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(HOST_APP + ".tmp", HOST_APP)
    return HOST_APP


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [os.path.join(CSRC, f) for f in os.listdir(CSRC)
            if f.endswith(".hip") or f.endswith(".cpp")]
    # the charged-particle and history kernels take longest: start them first
    order = ("_el_", "_split_", "_hist_", "_elw_", "_kerma_", "_ev_", "engine")
    srcs.sort(key=lambda f: next((i for i, k in enumerate(order) if k in f), len(order)))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    build_app()
    if verbose:
        print("built", LIB, "and", HOST_APP)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)

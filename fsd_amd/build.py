"""Build libfsd.so (HIP kernels for gfx950 + host C++) in-tree.

    python -m fsd_amd.build        # or fsd_amd.build.build()

Objects go to fsd_amd/_build/, the library to fsd_amd/libfsd.so.  hipcc
cross-compiles for gfx950 without a GPU present.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfsd.so")
ARCH = os.environ.get("FSD_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["k_pairs.hip", "k_mobility.hip", "k_precond.hip", "k_solver.hip", "k_brownian.hip", "k_jacobi.hip", "k_noalloc.hip", "k_learned.hip", "k_refine.hip"]
HOST_HIP_SOURCES = ["stokes.cpp", "autotune.cpp"]      # host code using the HIP runtime, compiled by hipcc
CXX_SOURCES = ["rcm.cpp", "sweep_plan.cpp", "precond_plan.cpp", "lanczos_host.cpp"]
C_SOURCES = ["ewald_table.c"]          # binary128 table (gcc + libquadmath)


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + p.stdout)
    return p.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    inc = ["-I" + CSRC, "-I" + os.path.join(ROOT, "include")]
    hip_flags = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH,
                 "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                 "-Wno-unused-result", "-Wno-unused-value", "-Wno-unknown-warning-option"] + inc
    jobs = []
    objs = []
    for s in HIP_SOURCES + HOST_HIP_SOURCES:
        o = os.path.join(BUILD, os.path.splitext(s)[0] + ".o")
        lang = ["-x", "hip"] if s.endswith(".cpp") else []
        jobs.append([hipcc] + hip_flags + lang + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in CXX_SOURCES:
        o = os.path.join(BUILD, os.path.splitext(s)[0] + ".o")
        jobs.append(["g++", "-O2", "-fPIC", "-std=c++17", "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in C_SOURCES:
        o = os.path.join(BUILD, os.path.splitext(s)[0] + ".o")
        jobs.append(["gcc", "-O2", "-fPIC", "-std=gnu11", "-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs +
         ["-L/opt/rocm/lib", "-lrocfft", "-lquadmath", "-lpthread", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

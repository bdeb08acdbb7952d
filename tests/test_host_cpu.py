"""CPU tests of the library's host side: its RCM and sweep schedules compiled
standalone with g++, the C-ABI surface of libfsd.so (symbols only, no GPU
calls), the no-GPU failure path, the workload generator and the bench
contract."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from _cases import packing

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "fsd_amd", "csrc")
CPU = os.path.join(ROOT, "tests", "cpu")


def _gxx(out, *srcs, shared=False):
    cmd = ["g++", "-O2", "-std=c++17", "-o", out] + (["-shared", "-fPIC"] if shared else []) + list(srcs)
    subprocess.run(cmd, check=True)


@pytest.fixture(scope="module")
def host_rcm(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("rcm") / "libhost_rcm.so")
    _gxx(so, os.path.join(CPU, "host_rcm_shim.cpp"), os.path.join(CSRC, "rcm.cpp"), shared=True)
    lib = C.CDLL(so)
    lib.test_host_rcm.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    return lib


def test_host_rcm_matches_reference_golden(host_rcm):
    """the library's RCM returns the reference GENRCM permutations"""
    z = np.load(os.path.join(ROOT, "tests", "golden", "rcm_golden.npz"))
    for name in z["names"]:
        phead = z[name + "_phead"].astype(np.int32)
        plist = z[name + "_plist"].astype(np.int32)
        N = len(phead) - 1
        perm = np.zeros(N, dtype=np.int32)
        host_rcm.test_host_rcm(N, phead.ctypes.data, plist.ctypes.data, perm.ctypes.data)
        assert np.array_equal(perm, z[name + "_perm"]), name


def test_host_rcm_matches_oracle(host_rcm):
    """library RCM == oracle RCM on pruned graphs of dense and dilute packings"""
    from oracle.oracle import Oracle
    from oracle import oracle as om
    for N, phi, seed in [(2000, 0.45, 4), (1500, 0.1, 9)]:
        pos, L = packing(N, phi, seed)
        o = Oracle(pos, L)
        _, phead, plist, _ = o.prune()
        a = np.zeros(N, dtype=np.int32)
        b = np.zeros(N, dtype=np.int32)
        host_rcm.test_host_rcm(N, phead.ctypes.data, plist.ctypes.data, a.ctypes.data)
        om.load().orc_rcm(N, phead.ctypes.data_as(C.c_void_p), plist.ctypes.data_as(C.c_void_p),
                          b.ctypes.data_as(C.c_void_p))
        assert np.array_equal(a, b)
        assert np.array_equal(np.sort(a), np.arange(N))


@pytest.mark.parametrize("args", [("1", "1", "5"), ("50", "5", "5"), ("3000", "40", "5"), ("20000", "200", "5"),
                                  ("3000", "30", "16"), ("20000", "300", "8", "30"), ("2000", "60", "70", "10"),
                                  ("40000", "400", "6", "50", "2048")])
def test_sweep_schedule_emulation(tmp_path, args):
    """tp_sweep_kernel's schedule replayed on the CPU with a visibility model:
    indices in bounds, ring and global reads find their producers after they
    are ordered, reductions stay in a wave, and the result equals a direct
    block solve (rows with up to 64 blocks, many isolated rows, the smallest
    ring the planner accepts)"""
    exe = str(tmp_path / "sweep_emulation")
    _gxx(exe, os.path.join(CPU, "sweep_emulation.cpp"), os.path.join(CSRC, "sweep_plan.cpp"))
    r = subprocess.run([exe] + list(args), capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "hazards=0 oob=0" in r.stdout


@pytest.mark.parametrize("N,seed", [(2000, 1), (5000, 7)])
def test_plan_keep_order(tmp_path, N, seed):
    """the host plan with the last ordering kept (precond mode 2 while the neighbour
    list stands): on a perturbed graph it keeps the ordering and builds the new graph's
    block pattern and level schedules under it; without keep_order it orders by the
    new graph's RCM"""
    exe = str(tmp_path / "plan_keep_order")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(CPU, "plan_keep_order.cpp"),
                    os.path.join(CSRC, "precond_plan.cpp"), os.path.join(CSRC, "rcm.cpp"),
                    os.path.join(CSRC, "sweep_plan.cpp"), "-lpthread"], check=True)
    r = subprocess.run([exe, str(N), str(seed)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad=0" in r.stdout


def test_tridiagonal_square_root(tmp_path):
    """FP64 T^(1/2) e_1 of the Lanczos sampler (lanczos_host.cpp) satisfies
    t^T t = T_11 and t^T T t = (T^2)_11 up to m = 99"""
    exe = str(tmp_path / "tridiag_sqrt_check")
    _gxx(exe, os.path.join(CPU, "tridiag_sqrt_check.cpp"), os.path.join(CSRC, "lanczos_host.cpp"))
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def _build_facade_example(out):
    cmd = ["g++", "-O2", "-std=c++17", "-Wno-unused-result", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + os.path.join(ROOT, "include"), os.path.join(CPU, "stokes_facade_example.cpp"),
           "-L" + os.path.join(ROOT, "fsd_amd"), "-lfsd", "-L/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath,/opt/rocm/lib", "-Wl,-rpath," + os.path.join(ROOT, "fsd_amd"), "-o", out]
    subprocess.run(cmd, check=True)


def test_cpp_facade_compiles(tmp_path):
    """the C++ Stokes facade (include/fsd_stokes.hpp) compiles and links against libfsd.so"""
    _build_facade_example(str(tmp_path / "stokes_facade_example"))


# ------------------------------------------------------------ library surface
def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "fsd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)
    return sorted(set(re.findall(r"\b(fsd_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header_symbols():
    """libfsd.so loads here (no GPU calls) and exports every function fsd.h declares;
    the ctypes binding declares the same set"""
    from fsd_amd import _lib
    lib = _lib.load()
    syms = _header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == syms


def test_no_gpu_fails_loudly():
    """without a GPU the library reports an error; nothing falls back to the CPU"""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from fsd_amd.stokes import Stokes
    from fsd_amd._lib import FsdError
    with pytest.raises((FsdError, RuntimeError)):
        Stokes(64, 12.0)


# ------------------------------------------------------------ workloads, bench
@pytest.mark.parametrize("N,phi,seed", [(343, 0.2, 11), (2000, 0.45, 4)])
def test_random_packing(N, phi, seed):
    """seeded packings: N spheres of radius 1 at volume fraction phi, no pair
    closer than 2 + 1e-3 under periodic minimum image"""
    from scipy.spatial import cKDTree
    pos, L = packing(N, phi, seed)
    assert pos.shape == (N, 3)
    assert N * 4.0 / 3.0 * np.pi / L ** 3 == pytest.approx(phi, rel=1e-6)
    t = cKDTree(np.mod(pos, L), boxsize=L)
    d, _ = t.query(np.mod(pos, L), k=2)
    assert d[:, 1].min() >= 2.0 + 1e-3 - 1e-6


def test_bench_rejects_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True)
    assert r.returncode == 2
    assert "one GPU" in r.stderr

"""GPU parity of the preconditioner applied by Jacobi iterations on its
triangular factors (precond modes 1 and 2) and of IC(0) computed by parallel
fixed-point sweeps, against the FP64 oracle."""
import numpy as np
import pytest

from tests._cases import packing, rel_err, rel_err2, forces, random_vec

pytestmark = pytest.mark.gpu

CASES = {"C1": (343, 0.20, 11), "C2": (4096, 0.30, 12)}


def _t(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32).reshape(-1)).cuda()


def _setup(case, **kw):
    import torch
    from fsd_amd.stokes import Stokes
    from oracle.oracle import Oracle
    N, phi, seed = CASES[case]
    pos, L = packing(N, phi, seed)
    print("case", case, "N", N, "options", kw, flush=True)
    st = Stokes(N, L, **kw)
    p4 = np.zeros((N, 4), dtype=np.float32)
    p4[:, :3] = pos
    dpos = torch.from_numpy(p4.reshape(-1)).cuda()
    st.build_nlist(dpos)
    orc = Oracle(pos, L)
    return st, orc, dpos, N


@pytest.mark.parametrize("case", ["C1", "C2"])
@pytest.mark.parametrize("k", [1, 2, 3])
def test_jacobi_apply_parity(case, k):
    """G v with k Jacobi iterations per triangular solve equals the oracle's
    k-step application of the same IC(0) factor (1e-4, the factor's tolerance)"""
    st, orc, dpos, N = _setup(case, precond="ichol_jacobi", jacobi_iters=k)
    st.build_precond(dpos)
    orc.build_precond()
    x = random_vec(6 * N, 4).astype(np.float32)
    y = st.precond_apply(_t(x)).cpu().numpy()
    e = rel_err(y, orc.precond_apply_jacobi(x, k))
    print("k", k, "G v rel err", e)
    assert e < 1e-4


@pytest.mark.parametrize("case", ["C1", "C2"])
def test_jacobi_deep_equals_exact_solves(case):
    """with more Jacobi iterations than the factor's DAG is deep the application is
    the exact G = P^T L^-T L^-1 P (oracle, 1e-4)"""
    st, orc, dpos, N = _setup(case, precond="ichol_jacobi", jacobi_iters=400)
    info = st.build_precond(dpos)
    orc.build_precond()
    assert info["levels"] < 400
    x = random_vec(6 * N, 5).astype(np.float32)
    e = rel_err(st.precond_apply(_t(x)).cpu().numpy(), orc.precond_apply(x))
    print("levels", info["levels"], "G v rel err", e)
    assert e < 1e-4


@pytest.mark.parametrize("case", ["C1", "C2"])
def test_ichol_parallel_sweeps_converge(case):
    """IC(0) by parallel fixed-point sweeps (Chow-Patel) reaches the factor of the
    level-scheduled recurrence: after as many sweeps as the DAG is deep it equals
    the oracle's IC(0) (1e-4)"""
    st, orc, dpos, N = _setup(case, precond="ichol_jacobi", factor_sweeps=1)
    info = st.build_precond(dpos)
    depth = info["levels"]
    st2, _, _, _ = _setup(case, precond="ichol_jacobi", factor_sweeps=depth + 1)
    st2.build_precond(dpos)
    orc.build_precond()
    L = st2.get_precond()["L"]
    e = rel_err(L, orc.L_val)
    e1 = rel_err(st.get_precond()["L"], orc.L_val)
    print("levels", depth, "one sweep rel err", e1, "converged rel err", e)
    assert e < 1e-4 and e1 > e


@pytest.mark.parametrize("case", ["C1", "C2"])
def test_solve_with_jacobi_preconditioner(case):
    """a saddle solve preconditioned with 3 Jacobi iterations (IC(0) from 3
    parallel sweeps) reaches the oracle's FP64 solution: true residual 1e-5
    checked with the FP64 operator, velocities within 1e-3 (what a 1e-5 residual
    bounds at these sizes)"""
    st, orc, dpos, N = _setup(case, precond="ichol_jacobi", jacobi_iters=3, factor_sweeps=3,
                              gmres_true_tol=1e-5, gmres_restart=100)
    st.build_precond(dpos)
    orc.build_precond()
    b = orc.rhs(forces(N, 22), 1.0)
    x, info = st.saddle_solve(dpos, _t(b))
    xg = x.cpu().numpy().astype(np.float64)
    r = np.linalg.norm(b - orc.saddle_multiply(xg)) / np.linalg.norm(b)
    xo, it, res = orc.gmres(b, tol=1e-11)
    e = rel_err2(xg[11 * N:], xo[11 * N:])
    print("iterations", info["iterations"], "FP64 residual", r, "velocity rel err", e)
    assert r < 1.5e-5
    assert e < 1e-3

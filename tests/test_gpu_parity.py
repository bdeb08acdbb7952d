"""GPU parity of every operator of the FSD step against the FP64 oracle.

Integer and index results (neighbour sets, pruned lists, RCM permutation,
sparsity patterns) must match bit-exactly; floating-point results within the
tolerance written in each test (FP32 kernels vs an FP64 restatement of the
same discretisation; the velocity bar is 1e-5 relative).
"""
import numpy as np
import pytest

from tests._cases import packing, rel_err, rel_err2, forces, random_vec, sorted_lists

pytestmark = pytest.mark.gpu

CASES = {
    # C0: a box of fewer than three neighbour-list cells per side (list built over all
    # pairs, no cell order for the real-space and lubrication gathers)
    "C0": (64, 0.14, 10),
    "C1": (343, 0.20, 11),
    "C2": (4096, 0.30, 12),
}


def _setup(case, **kw):
    import torch
    from fsd_amd.stokes import Stokes
    from oracle.oracle import Oracle
    N, phi, seed = CASES[case]
    pos, L = packing(N, phi, seed)
    print("case", case, "N", N, "phi", phi, "seed", seed, flush=True)
    st = Stokes(N, L, **kw)
    p4 = np.zeros((N, 4), dtype=np.float32)
    p4[:, :3] = pos
    dpos = torch.from_numpy(p4.reshape(-1)).cuda()
    st.build_nlist(dpos)
    orc = Oracle(pos, L)
    return st, orc, dpos, N


def _t(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()


@pytest.mark.parametrize("case", ["C0", "C1", "C2"])
def test_nlist_bitexact(case):
    st, orc, dpos, N = _setup(case)
    nn, nb = st.get_nlist()
    got = sorted_lists(nn, nb)
    ref = sorted_lists(orc.nneigh, orc.nbrs)
    assert got == ref


@pytest.mark.parametrize("case", ["C0", "C1", "C2"])
def test_rfu_parity(case):
    st, orc, dpos, N = _setup(case)
    if case == "C0":   # the case is meant to exercise the all-pairs list
        L = packing(*CASES[case])[1]
        assert L / (st.pse_info()["rcut"] + 0.4) < 3
    U = random_vec(6 * N, 1)
    F = st.rfu(dpos, _t(U)).cpu().numpy()
    Fo = orc.rfu(U.astype(np.float32))
    e = rel_err(F, Fo)
    print("RFU rel err", e)
    assert e < 1e-5


@pytest.mark.parametrize("case", ["C0", "C1", "C2"])
def test_rfe_rsu_rse_parity(case):
    st, orc, dpos, N = _setup(case)
    F = st.rfe(dpos, 1.0).cpu().numpy()
    assert rel_err(F, orc.rfe(1.0)) < 1e-5
    U = random_vec(6 * N, 2).astype(np.float32)
    S = st.rsu(dpos, _t(U)).cpu().numpy()
    assert rel_err(S, orc.rsu(U)) < 1e-5
    S = st.rse(dpos, 1.0).cpu().numpy()
    assert rel_err(S, orc.rse(1.0)) < 1e-5


@pytest.mark.parametrize("case", ["C0", "C1", "C2"])
@pytest.mark.parametrize("parts", [1, 2, 3])
def test_mobility_parity(case, parts):
    st, orc, dpos, N = _setup(case)
    gF = random_vec(11 * N, 3).astype(np.float32)
    gU = st.mobility(dpos, _t(gF), parts).cpu().numpy()
    gUo = orc.mobility(gF, parts)
    e = rel_err(gU, gUo)
    print("mobility parts", parts, "rel err", e, "l2", rel_err2(gU, gUo))
    assert e < 1e-5


@pytest.mark.parametrize("N,phi,seed", [(4096, 0.30, 12), (32768, 0.45, 1)])
def test_fused_x_transform_matches_3d_plans(N, phi, seed, monkeypatch):
    """on grids with Nx = 8 B (40 and 72 here), the (y, z) transforms by rocFFT plus the
    kernel doing x and the Green's function in one pass give the mobility of the 3-D
    rocFFT plans and the separate Green's function kernel (FP32, different rounding)"""
    import torch
    from fsd_amd.stokes import Stokes
    pos, L = packing(N, phi, seed)
    print("N", N, "phi", phi, "seed", seed, flush=True)
    p4 = np.zeros((N, 4), dtype=np.float32)
    p4[:, :3] = pos
    dpos = torch.from_numpy(p4.reshape(-1)).cuda()
    gF = random_vec(11 * N, 3).astype(np.float32)
    out = {}
    for f3 in ("0", "1"):
        monkeypatch.setenv("FSD_FFT_3D", f3)
        st = Stokes(N, L)
        st.build_nlist(dpos)
        assert st.pse_info()["Nx"] % 8 == 0
        out[f3] = [st.mobility(dpos, _t(gF), parts).cpu().numpy() for parts in (2, 3)]
    for k, parts in enumerate((2, 3)):
        e = rel_err(out["0"][k], out["1"][k])
        print("parts", parts, "fused vs 3-D rel err", e)
        assert e < 1e-5


@pytest.mark.parametrize("case", ["C1", "C2"])
def test_precond_pattern_bitexact(case):
    st, orc, dpos, N = _setup(case)
    info = st.build_precond(dpos)
    g = st.get_precond()
    orc.build_precond()
    assert np.array_equal(g["pn"], orc.pn)
    assert np.array_equal(g["plist"], orc.plist)
    assert np.array_equal(g["perm"], orc.perm)
    assert np.array_equal(g["rowptr"], orc.A_rowptr)
    assert np.array_equal(g["col"], orc.A_col)
    assert np.array_equal(g["prow"], orc.B_rowptr)
    assert np.array_equal(g["pcol"], orc.B_col)
    assert info["nnz"] == len(orc.A_col)
    assert info["ichol_relaxer"] == orc.lam
    ev = rel_err(g["val"], orc.A_val)
    eL = rel_err(g["L"], orc.L_val)
    print("A values rel err", ev, "L values rel err", eL, "lambda", orc.lam)
    assert ev < 1e-5
    assert eL < 1e-4


@pytest.mark.parametrize("case", ["C1", "C2"])
def test_precond_apply_parity(case):
    st, orc, dpos, N = _setup(case)
    st.build_precond(dpos)
    orc.build_precond()
    x = random_vec(6 * N, 4).astype(np.float32)
    y = st.precond_apply(_t(x)).cpu().numpy()
    yo = orc.precond_apply(x)
    e = rel_err(y, yo)
    print("G v rel err", e)
    assert e < 1e-4


@pytest.mark.parametrize("case", ["C1", "C2"])
def test_saddle_operators_parity(case):
    st, orc, dpos, N = _setup(case)
    st.build_precond(dpos)
    orc.build_precond()
    x = random_vec(17 * N, 5).astype(np.float32)
    y = st.saddle_multiply(dpos, _t(x)).cpu().numpy()
    yo = orc.saddle_multiply(x)
    e1 = rel_err(y, yo)
    z = st.saddle_precond(_t(x)).cpu().numpy()
    zo = orc.saddle_precond(x)
    e2 = rel_err(z, zo)
    print("saddle multiply rel err", e1, "saddle precond rel err", e2)
    assert e1 < 1e-5
    assert e2 < 1e-4


def test_solve_C1_velocities():
    """one C1 velocity solve with the reference's GMRES (left preconditioned,
    |s| <= 1e-6 ||b||) against the oracle solved to 1e-11"""
    st, orc, dpos, N = _setup("C1", gmres_tol=1e-6)
    st.build_precond(dpos)
    orc.build_precond()
    F = forces(N, 21)
    b = orc.rhs(F, 0.0)
    x, info = st.saddle_solve(dpos, _t(b))
    xo, it, res = orc.gmres(b, tol=1e-11)
    U, Uo = x.cpu().numpy()[11 * N:], xo[11 * N:]
    e = rel_err2(U, Uo)
    print("GMRES iters", info["iterations"], "true residual", info["true_residual"], "velocity rel err", e,
          "oracle iters", it)
    assert info["converged"]
    assert e < 1e-5


@pytest.mark.parametrize("case", ["C1", "C2"])
def test_solve_true_residual_mode(case):
    """gmres_true_tol: the right-preconditioned solve stops at a true relative
    residual of 1e-4, checked independently with the oracle's FP64 operator"""
    st, orc, dpos, N = _setup(case, gmres_true_tol=1e-4)
    st.build_precond(dpos)
    orc.build_precond()
    F = forces(N, 22)
    b = orc.rhs(F, 1.0)
    x, info = st.saddle_solve(dpos, _t(b))
    xg = x.cpu().numpy().astype(np.float64)
    r = np.linalg.norm(b - orc.saddle_multiply(xg)) / np.linalg.norm(b)
    xo, it, res = orc.gmres(b, tol=1e-10)
    e = rel_err2(xg[11 * N:], xo[11 * N:])
    print("iters", info["iterations"], "reported", info["true_residual"], "FP64 residual", r, "velocity err", e)
    assert info["converged"] and info["true_residual"] <= 1e-4
    assert r < 1.2e-4
    assert e < 1e-2             # a 1e-4 residual bounds the velocity error only up to the conditioning


@pytest.mark.parametrize("case", ["C1", "C2"])
def test_solve_one_gram_schmidt_pass(case):
    """one classical Gram-Schmidt pass per Arnoldi step (gs_passes = 1) reaches the
    same true residual, checked with the oracle's FP64 operator, and the same
    solution as two passes to within that residual"""
    import torch
    from fsd_amd.stokes import Stokes
    st2, orc, dpos, N = _setup(case, gmres_true_tol=1e-5)
    orc.build_precond()
    b = orc.rhs(forces(N, 24), 1.0)
    xs = []
    for passes in (2, 1):
        st = st2 if passes == 2 else Stokes(N, packing(*CASES[case])[1], gmres_true_tol=1e-5, gram_schmidt_passes=1)
        if passes == 1:
            st.build_nlist(dpos)
        st.build_precond(dpos)
        x, info = st.saddle_solve(dpos, _t(b))
        xg = x.cpu().numpy().astype(np.float64)
        r = np.linalg.norm(b - orc.saddle_multiply(xg)) / np.linalg.norm(b)
        print("passes", passes, "iters", info["iterations"], "FP64 residual", r)
        assert info["converged"] and r < 3e-5
        xs.append(xg)
    e = rel_err2(xs[1][11 * N:], xs[0][11 * N:])
    print("velocity rel diff", e)
    assert e < 1e-3


def test_step_C1():
    """fsd_step with injected forces: velocities equal the oracle's solve and
    positions move by (U + shear) dt"""
    import torch
    from fsd_amd.stokes import ParticleState
    st, orc, dpos, N = _setup("C1", gmres_tol=1e-6, dt=1e-3)
    pos, L = packing(*CASES["C1"])
    state = ParticleState(pos)
    F = forces(N, 22)
    info = st.integrateStepOne(0, state, force6=_t(F), shear_rate=0.0)
    vel = state.velocity.cpu().numpy()
    orc.build_precond()
    b = orc.rhs(F, 0.0)
    xo, it, res = orc.gmres(b, tol=1e-11)
    e = rel_err2(vel[:6 * N], xo[11 * N:])
    print("step iterations", info["iterations"], "velocity rel err", e)
    assert e < 1e-5
    newpos = state.positions()
    disp = newpos - pos
    disp -= L * np.rint(disp / L)
    expect = vel[:6 * N].reshape(N, 6)[:, :3] * 1e-3
    assert np.abs(disp - expect).max() < 1e-5


def test_nlist_kept_until_half_buffer_moved():
    """the step keeps its neighbour list while no particle has moved r_buff / 2
    since it was built (HOOMD's distance check) and rebuilds it after; a kept
    list gives the velocities of a freshly built one (FP32 summation order
    aside), over steps under shear"""
    import torch
    from fsd_amd.stokes import Stokes, ParticleState
    N, phi, seed = CASES["C1"]
    pos, L = packing(N, phi, seed)
    F = _t(forces(N, 23))
    # A keeps its list; B rebuilds it before every step (an explicit build)
    a, b = Stokes(N, L, gmres_tol=1e-7, dt=0.01), Stokes(N, L, gmres_tol=1e-7, dt=0.01)
    sa, sb = ParticleState(pos), ParticleState(pos)
    rebuilt, worst = [], 0.0
    for k in range(8):
        b.build_nlist(sb.pos)
        ia = a.integrateStepOne(k, sa, force6=F, shear_rate=0.5)
        ib = b.integrateStepOne(k, sb, force6=F, shear_rate=0.5)
        rebuilt.append(ia["nlist_rebuilt"])
        assert ib["nlist_rebuilt"] == 0
        va, vb = sa.velocity.cpu().numpy(), sb.velocity.cpu().numpy()
        worst = max(worst, rel_err2(va, vb))
        sb.pos.copy_(sa.pos)   # same positions for the next step
        sb.image.copy_(sa.image)
    print("rebuilt", rebuilt, "worst velocity rel diff", worst)
    assert rebuilt[0] == 1 and 0 in rebuilt[1:] and 1 in rebuilt[1:]
    assert worst < 1e-5


def test_force_law_parity():
    st, orc, dpos, N = _setup("C1")
    F = st.force_law(dpos, T_ext=0.3).cpu().numpy()
    Fo = orc.force_law(T_ext=0.3)
    assert rel_err(F, Fo) < 1e-5


def test_psev3_surface_matches_stokes():
    """integrate.PSEv3 over fsd_amd.system runs the same steps as driving
    Stokes directly (steady shear 1, reference force law)"""
    import torch
    from fsd_amd import system, shear_function, integrate
    from fsd_amd.stokes import Stokes, ParticleState
    N, phi, seed = CASES["C1"]
    pos, L = packing(N, phi, seed)
    sysm = system.System(pos, L, dt=1e-3)
    igr = integrate.PSEv3(sysm.group_all(), T=0.0, function_form=shear_function.steady(1e-3, 1.0), gmres_tol=1e-6)
    info = sysm.run(2)
    st = Stokes(N, L, gmres_tol=1e-6, dt=1e-3)
    state = ParticleState(pos)
    for k in range(2):
        st.integrateStepOne(k, state, shear_rate=1.0)
    a, b = sysm.positions(), state.positions()
    print("PSEv3 info", info, "max |dx|", np.abs(a - b).max())
    assert info["converged"]
    assert np.allclose(a, b, atol=1e-6)
    assert np.abs(a - pos).max() > 0


def test_cpp_facade_runs(tmp_path):
    """the C++ Stokes facade drives two sheared steps through the C ABI"""
    import subprocess
    import os
    from tests.test_host_cpu import _build_facade_example
    exe = str(tmp_path / "stokes_facade_example")
    _build_facade_example(exe)
    table = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fsd_amd", "data",
                         "resistance_table.txt")
    r = subprocess.run([exe, table], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0


def test_friction_type1():
    """setFriction TYPE1: the library's table equals the oracle's bit for bit,
    and R_FU with friction matches the oracle"""
    import ctypes as C
    from fsd_amd.tables import load_resistance_table
    from oracle.oracle import Oracle, friction_type1
    st, orc, dpos, N = _setup("C1")
    h0, alpha = 0.01, [0.5, 0.4, 0.3, 0.2, 0.1, 0.05]
    st.setFriction("TYPE1", h0, alpha)
    got = np.zeros(22000, dtype=np.float32)
    assert st.lib.fsd_get_resistance_table(st.ctx, got.ctypes.data_as(C.c_void_p)) == 0
    dist, vals = load_resistance_table()
    want = friction_type1(dist, vals, h0, alpha, parity=True)
    assert np.array_equal(got, want)
    pos, L = packing(*CASES["C1"])
    o2 = Oracle(pos, L, res_table=(dist, want))
    U = random_vec(6 * N, 9).astype(np.float32)
    e = rel_err(st.rfu(dpos, _t(U)).cpu().numpy(), o2.rfu(U))
    print("RFU with friction rel err", e)
    assert e < 1e-5


def test_box_resize_shear_variant():
    """box deformation with a wrapped strain (variant.shear_variant): every step
    converges and the particles stay inside the tilted box"""
    from fsd_amd import system, shear_function, integrate, variant
    N, phi, seed = CASES["C1"]
    pos, L = packing(N, phi, seed)
    sysm = system.System(pos, L, dt=1e-2)
    f = shear_function.steady(1e-2, 20.0)
    sysm.box_resize(xy=variant.shear_variant(f, total_timestep=100, max_strain=0.5))
    integrate.PSEv3(sysm.group_all(), T=0.0, function_form=f, gmres_tol=1e-5)
    tilts = []
    for _ in range(4):
        info = sysm.run(1)
        tilts.append(sysm.box[3])
        assert info["converged"]
    p = sysm.positions()
    xy = sysm.box[3]
    print("tilts", tilts)
    assert min(tilts) < 0 < max(tilts)                       # the strain wrapped
    assert np.all(np.abs(p[:, 1]) <= L / 2 + 1e-4) and np.all(np.abs(p[:, 2]) <= L / 2 + 1e-4)
    assert np.all(np.abs(p[:, 0] - xy * p[:, 1]) <= L / 2 + 1e-4)


@pytest.mark.parametrize("N,phi,seed", [(4096, 0.30, 12), (8192, 0.45, 3)])
def test_precond_one_launch_matches_level_path(N, phi, seed, monkeypatch):
    """the single-workgroup sweeps (schedule of sweep_plan.cpp, LDS ring,
    folded diagonal blocks; one workgroup per right-hand side by default, or
    both in one) give G v of the one-launch-per-level sweeps, for one and two
    right-hand sides (FP32, different rounding order)"""
    import torch
    from fsd_amd.stokes import Stokes
    pos, L = packing(N, phi, seed)
    print("N", N, "phi", phi, "seed", seed, flush=True)
    p4 = np.zeros((N, 4), dtype=np.float32)
    p4[:, :3] = pos
    dpos = torch.from_numpy(p4.reshape(-1)).cuda()
    sts, infos = [], []
    for levels, joint in (("0", "0"), ("0", "1"), ("1", "0")):
        monkeypatch.setenv("FSD_TRSV_LEVELS", levels)
        monkeypatch.setenv("FSD_SWEEP_JOINT", joint)
        st = Stokes(N, L)
        st.build_nlist(dpos)
        infos.append(st.build_precond(dpos))
        sts.append(st)
    print("one-launch", infos[0], "level path", infos[2])
    assert infos[0]["sweep_steps"] > infos[0]["levels"] and infos[2]["sweep_steps"] == 0
    x = random_vec(6 * N, 5).astype(np.float32)
    y = [s.precond_apply(_t(x)).cpu().numpy() for s in sts]
    v = random_vec(17 * N, 6).astype(np.float32)
    z = [s.saddle_precond(_t(v)).cpu().numpy() for s in sts]
    for k in (0, 1):
        e1, e2 = rel_err(y[k], y[2]), rel_err(z[k], z[2])
        print("variant", k, "G v vs levels rel err", e1, "saddle precond", e2)
        assert e1 < 1e-5 and e2 < 1e-5


def test_rk2_stages_match_oracle():
    """the two Runge-Kutta stages (Integrator.cu:232-438) through the C-ABI equal the
    oracle's FP64 restatement, with shear, squirmer slip B1, angular noise and a wrap"""
    import torch
    from fsd_amd.stokes import Stokes
    from oracle.oracle import rk2_stage1, rk2_stage2
    N, L, dt, g, B1 = 343, 19.3, 0.05, 0.7, 0.4
    rng = np.random.default_rng(31)
    pos = np.zeros((N, 4), np.float32)
    pos[:, :3] = rng.uniform(-L / 2, L / 2, (N, 3))
    pos[:, 3] = 1.0
    ori = rng.standard_normal((N, 3)).astype(np.float32)
    ori /= np.linalg.norm(ori, axis=1, keepdims=True)
    gb = pos.copy()
    img = np.zeros((N, 3), np.int32)
    u = rng.uniform(-20, 20, (N, 6)).astype(np.float32)
    nz = rng.standard_normal((N, 3)).astype(np.float32)
    st = Stokes(N, L, sqm_B1=B1)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a).reshape(-1)).cuda()
    dgb = d(gb)
    po, oo = st.rk2_stage1(d(pos), d(ori), dgb, d(u), dt, g, noise3=d(nz))
    p1, o1, g1 = rk2_stage1(pos, ori, gb, u.astype(np.float64), dt, g, B1=B1, noise3=nz)
    po, oo, gbg = po.cpu().numpy().reshape(N, 4), oo.cpu().numpy().reshape(N, 3), dgb.cpu().numpy().reshape(N, 4)
    e1 = max(np.abs(po - p1).max(), np.abs(oo - o1).max(), np.abs(gbg - g1).max())
    dimg = d(img)
    pq, oq = st.rk2_stage2(0.5, d(pos), d(ori), 0.5, d(p1), d(o1), d(g1), d(u), dimg, 0.5, dt, g, noise3=d(nz))
    p2, o2, g2, i2 = rk2_stage2(0.5, pos, ori, 0.5, p1, o1, g1, u.astype(np.float64), img, L, 0.5, dt, g, B1=B1,
                                noise3=nz)
    pq, oq = pq.cpu().numpy().reshape(N, 4), oq.cpu().numpy().reshape(N, 3)
    e2 = max(np.abs(pq - p2).max(), np.abs(oq - o2).max())
    print("stage 1 max abs err", e1, "stage 2", e2)
    assert e1 < 2e-5 and e2 < 2e-5
    assert np.array_equal(dimg.cpu().numpy().reshape(N, 3), i2) and np.any(i2 != 0)

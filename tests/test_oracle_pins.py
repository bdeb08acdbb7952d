"""CPU pins of the FP64 oracle: closed forms, invariants, the reference's own
resistance table and RCM, and golden vectors.  Each test is chosen so that a
plausible mistake in the oracle (dropped term, wrong sign or index, transposed
operand) fails it."""
import ctypes as C
import os

import numpy as np
import pytest

from _cases import packing

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _oracle(*a, **kw):
    from oracle.oracle import Oracle
    return Oracle(*a, **kw)


def _grand(o, N, cols, parts=3):
    """columns `cols` of the 11N x 11N grand mobility"""
    out = np.zeros((11 * N, len(cols)))
    for k, c in enumerate(cols):
        g = np.zeros(11 * N)
        g[c] = 1.0
        out[:, k] = o.mobility(g, parts)
    return out


# ------------------------------------------------------------ resistance table
def test_resistance_table_anchors():
    """Stokes_ResistanceTable.cc: 1000 rows, h = r - 2 log-spaced from 1e-4 with
    step 0.004305 in log10; the anchors the lookup's clamps rely on"""
    from fsd_amd.tables import load_resistance_table, CONTACT_ROW
    dist, vals = load_resistance_table()
    v = vals.reshape(1000, 22)
    assert dist.shape == (1000,) and vals.shape == (22000,)
    assert dist[0] == np.float32(2.0001)
    assert v[0, 0] == np.float32(2500.839791) and v[0, 1] == np.float32(-2500.837257)
    assert CONTACT_ROW == 232 and dist[232] == np.float32(2.000997)
    assert dist[999] == np.float32(4.0) and np.all(v[999] == 0.0)
    # log10 spacing whose rounded step 0.004305 is the lookup's divisor
    h = dist[:999].astype(np.float64) - 2.0
    assert np.all(np.diff(h) > 0)
    assert abs(np.log10(h[998] / h[0]) / 998 - 0.004305) < 1e-6
    # pair and self squeeze terms diverge together: XA11 + XA12 stays O(1) at contact
    assert abs(v[0, 0] + v[0, 1]) < 0.01 and v[0, 0] > 2000


def test_rfu_two_sphere_table_values():
    """R_FU of two spheres along x at a table node returns the table's XA/YA/XC/YC
    entries (Lubrication.cu:Lubrication_RFU_kernel); along y the roles swap"""
    from fsd_amd.tables import load_resistance_table
    dist, vals = load_resistance_table()
    v = vals.reshape(1000, 22)
    row = 500
    r = float(dist[row])
    L = 30.0
    for axis in (0, 1):
        pos = np.zeros((2, 3))
        pos[1, axis] = r
        o = _oracle(pos, L)
        for comp, (c11, c12) in {axis: (0, 1), (axis + 1) % 3: (2, 3), 3 + axis: (6, 7),
                                 3 + (axis + 1) % 3: (8, 9)}.items():
            U = np.zeros(12)
            U[comp] = 1.0
            F = o.rfu(U)
            assert F[comp] == pytest.approx(v[row, c11], rel=2e-3)
            assert F[6 + comp] == pytest.approx(v[row, c12], rel=2e-3)


def test_rfu_symmetric():
    """R_FU is symmetric: x . R y == y . R x on a C1 packing"""
    pos, L = packing(343, 0.2, 11)
    o = _oracle(pos, L)
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal((2, 6 * 343))
    a, b = x @ o.rfu(y), y @ o.rfu(x)
    assert abs(a - b) <= 1e-10 * abs(a)


# ------------------------------------------------------------ mobility (PSE)
@pytest.mark.parametrize("L", [20.0, 30.0])
def test_hasimoto_self_mobility(L):
    """one sphere in a periodic cube (the pi typo off): translational self-mobility
    1 - 2.837297/L + 4.18879/L^3 (Hasimoto 1959; units 6 pi eta a = 1), rotational
    0.75 (1 - (4 pi / 3)/L^3) (the periodic rotlet's -T/(6 eta L^3))"""
    o = _oracle(np.zeros((1, 3)), L, error=1e-6, parity=False)
    M = _grand(o, 1, [0, 3])
    has = 1.0 - 2.837297 / L + 4.18879 / L ** 3
    assert M[0, 0] == pytest.approx(has, rel=1e-6)
    assert M[1, 0] == pytest.approx(0.0, abs=1e-9) and M[2, 0] == pytest.approx(0.0, abs=1e-9)
    assert M[3, 1] == pytest.approx(0.75 * (1.0 - 4.18879 / L ** 3), rel=2e-6)


def test_rpy_pair_limit():
    """pair blocks in a large box: free-space RPY (overlapping form for r < 2)
    plus the leading periodic shift -2.837297/L; the next correction is O(r^2/L^3)"""
    L = 80.0
    pos = np.array([[0.0, 0.0, 0.0], [1.8, 2.4, 0.0], [-0.9, 0.0, 1.2]])   # r = 3.0 and r = 1.5 from particle 0
    o = _oracle(pos, L, error=1e-6, parity=False)
    M = _grand(o, 3, [0, 1, 2])
    for j in (1, 2):
        rv = pos[j] - pos[0]
        r = np.linalg.norm(rv)
        e = rv / r
        if r >= 2:
            R = (3 / (4 * r) + 1 / (2 * r ** 3)) * np.eye(3) + (3 / (4 * r) - 3 / (2 * r ** 3)) * np.outer(e, e)
        else:
            R = (1 - 9 * r / 32) * np.eye(3) + (3 * r / 32) * np.outer(e, e)
        blk = M[6 * j:6 * j + 3, :]
        assert np.abs(blk - (R - 2.837297 / L * np.eye(3))).max() < 1e-4, (r, blk, R)


def test_ewald_split_independence():
    """the grand mobility does not depend on the Ewald splitting xi for pairs
    inside every xi's real-space cutoff (pairs beyond a cutoff carry the
    truncation error the reference's rcut rule allows)"""
    pos = np.array([[0.0, 0.0, 0.0], [1.5, 2.0, 0.0], [0.0, -2.2, 1.0]])
    L = 20.0
    Ms = [_grand(_oracle(pos, L, xi=xi, error=1e-6, parity=False), 3, range(33)) for xi in (0.5, 0.6, 0.8)]
    assert np.abs(Ms[1] - Ms[0]).max() < 5e-6
    assert np.abs(Ms[2] - Ms[0]).max() < 5e-6


def test_mobility_symmetric():
    """the grand mobility is symmetric on a random dilute configuration"""
    pos, L = packing(60, 0.05, 5)
    o = _oracle(pos, L, error=1e-6, parity=False)
    rng = np.random.default_rng(1)
    x, y = rng.standard_normal((2, 11 * 60))
    a, b = x @ o.mobility(y), y @ o.mobility(x)
    assert abs(a - b) <= 1e-7 * abs(a)


def test_ewald_table_matches_functions():
    """the float table nodes r_k = (k+1) dr equal the closed-form functions
    (fsd_ewald_functions) rounded to float (Stokes.cc: Ewald real-space table)"""
    from oracle import oracle as om
    xi = 0.5
    nR = 2000
    tab = om.ewald_table(xi, nR).reshape(nR, 8)
    for k in (0, 499, 1999):
        r = (k + 1) * 0.001
        f = om.ewald_functions(r, xi)
        assert tab[k, 0] == pytest.approx(f[0], rel=1e-6, abs=1e-7)
        assert tab[k, 1] == pytest.approx(f[1], rel=1e-6, abs=1e-7)


# ------------------------------------------------------------ preconditioner
def test_ichol_reproduces_matrix_on_pattern():
    """IC(0): (L L^T)_ij = B_ij + lambda * shift_ii on the stored lower pattern
    (csric02 on the RCM-permuted R_FU^nf with the diagonal relaxer)"""
    pos, L = packing(343, 0.2, 11)
    o = _oracle(pos, L)
    o.build_precond()
    n = 6 * 343
    rp, col, bval, lval = o.B_rowptr, o.B_col, o.B_val, o.L_val
    Lm = np.zeros((n, n))
    Bm = np.zeros((n, n))
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            j = col[k]
            Bm[i, j] = bval[k]
            if j <= i:
                Lm[i, j] = lval[k]
    shift = np.tile([1.0, 1.0, 1.0, 1.33333333, 1.33333333, 1.33333333], 343) * o.lam
    LL = Lm @ Lm.T
    mask = np.tril(Bm != 0)
    target = Bm + np.diag(shift)
    assert np.abs((LL - target)[mask]).max() < 1e-9 * np.abs(target).max()


def _ref_rcm():
    path = os.path.join(ROOT, "oracle", "_ref", "libref_rcm.so")
    if not os.path.exists(path):
        return None
    lib = C.CDLL(path)
    lib.ref_genrcm.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    return lib


def _oracle_rcm(phead, plist):
    from oracle import oracle as om
    N = len(phead) - 1
    perm = np.zeros(N, dtype=np.int32)
    om.load().orc_rcm(N, phead.ctypes.data_as(C.c_void_p), plist.ctypes.data_as(C.c_void_p),
                      perm.ctypes.data_as(C.c_void_p))
    return perm


def test_rcm_golden_vectors():
    """oracle RCM equals the reference GENRCM's permutations stored in
    tests/golden/rcm_golden.npz (written by tests/golden/make_rcm_golden.py)"""
    z = np.load(os.path.join(GOLDEN, "rcm_golden.npz"))
    for name in z["names"]:
        phead, plist, perm = z[name + "_phead"], z[name + "_plist"], z[name + "_perm"]
        got = _oracle_rcm(phead.astype(np.int32), plist.astype(np.int32))
        assert np.array_equal(got, perm), name


def test_rcm_against_reference_build():
    """oracle RCM equals the reference GENRCM (oracle/_ref) on pruned graphs of
    packings, including isolated particles and several components"""
    lib = _ref_rcm()
    if lib is None:
        pytest.skip("oracle/_ref/libref_rcm.so not built (reference tree not mounted)")
    for N, phi, seed in [(343, 0.2, 11), (500, 0.05, 3), (1000, 0.35, 7)]:
        pos, L = packing(N, phi, seed)
        o = _oracle(pos, L)
        pn, phead, plist, _ = o.prune()
        ref = np.zeros(N, dtype=np.int32)
        lib.ref_genrcm(N, int(phead[N]), phead.ctypes.data, plist.ctypes.data, ref.ctypes.data)
        assert np.array_equal(_oracle_rcm(phead, plist), ref), (N, phi)


def test_friction_type1_table():
    """TYPE1 friction (Stokes.cc:1132-1191): the added term
    alpha (2 h^2/h0^3 - 3 h/h0^2 + 1/h) vanishes at h = h0, rows with h > h0 and
    all other columns are untouched, and parity moves the six columns"""
    from fsd_amd.tables import load_resistance_table
    from oracle.oracle import friction_type1
    dist, vals = load_resistance_table()
    k = 400
    h0 = float(np.float32(float(dist[k]) - 2.0))
    alpha = [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]
    for parity, cols in [(True, [3, 4, 5, 6, 9, 10]), (False, [2, 3, 4, 5, 8, 9])]:
        v = friction_type1(dist, vals, h0, alpha, parity).reshape(1000, 22)
        v0 = vals.reshape(1000, 22)
        others = [c for c in range(22) if c not in cols]
        assert np.array_equal(v[:, others], v0[:, others])
        assert np.array_equal(v[k + 1:], v0[k + 1:])
        assert np.allclose(v[k], v0[k], rtol=0, atol=1e-3 * np.abs(v0[k]).max())
        d = v[:k, cols] - v0[:k, cols]
        # sign pattern (+, -, -, +, +, +) with alpha > 0 and a positive polynomial below h0
        assert np.all(d[:, [0, 3, 4, 5]] > 0) and np.all(d[:, [1, 2]] < 0)


# ------------------------------------------------- R_FE, R_SU, R_SE, force law
# Table columns (Stokes_ResistanceTable.cc): the two-sphere functions of
# Jeffrey (1992) and Kim & Karrila (1991, ch. 11).  The pins below place two
# spheres on a coordinate axis at a table node and apply one velocity or one
# strain mode, for which the force, torque or stresslet of sphere 0 is a single
# table entry (or the sum of its "11" and "12" entries).
COL = {n: k for k, n in enumerate("XA11 XA12 YA11 YA12 YB11 YB12 XC11 XC12 YC11 YC12 XG11 XG12 YG11 YG12 "
                                   "YH11 YH12 XM11 XM12 YM11 YM12 ZM11 ZM12".split())}
ROW = 500          # a node well inside the table (h ~ 0.14)


def _two_spheres(axis):
    from fsd_amd.tables import load_resistance_table
    dist, vals = load_resistance_table()
    v = vals.reshape(1000, 22).astype(np.float64)
    pos = np.zeros((2, 3))
    pos[1, axis] = float(dist[ROW])
    return _oracle(pos, 30.0), (lambda name: v[ROW, COL[name]])


def _S5_to_tensor(s):
    """stresslet 5-vector [Sxx, Sxy, Sxz, Syz, Syy] -> traceless 3x3"""
    xx, xy, xz, yz, yy = s
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, -xx - yy]])


def _E_tensors(shear_rate, B2, mask2, ori):
    """per-particle strain tensors of the oracle's R_FE / R_SE (Lubrication.cu:389-418):
    E_inf with 2 Exy = shear rate, plus 0.6 B2 mask (p p^T - |p|^2 I/3), the
    traceless part (the float director is unit only to rounding)"""
    out = []
    for m, p in zip(mask2, ori):
        p = np.asarray(p, dtype=np.float64)
        E = np.zeros((3, 3))
        E[0, 1] = E[1, 0] = 0.5 * shear_rate
        E += 0.6 * B2 * float(m) * (np.outer(p, p) - (p @ p) * np.eye(3) / 3.0)
        out.append(E)
    return out


def test_rfe_two_sphere_table_columns():
    """R_FE:E for two spheres on the x axis (r_hat from 0 to 1 = +x): simple shear
    2Exy = 1 gives sphere 0 the force -(YG11 - YG12) y and torque -(YH11 + YH12) z
    (the transverse and rotational coupling functions); uniaxial extension
    E = e (x x - I/3) gives -(2e/3)(XG11 - XG12) x.  Sphere 1 gets the opposite force
    and the same torque (Jeffrey 1992: G21 = -G12, H21 = H12)."""
    o, t = _two_spheres(0)
    F = o.rfe(1.0)
    assert F[1] == pytest.approx(-(t("YG11") - t("YG12")), rel=1e-4)
    assert F[5] == pytest.approx(-(t("YH11") + t("YH12")), rel=1e-4)
    assert abs(F[0]) + abs(F[2]) + abs(F[3]) + abs(F[4]) < 1e-9 * abs(F[1])
    assert F[7] == pytest.approx(-F[1], rel=1e-9) and F[11] == pytest.approx(F[5], rel=1e-9)
    B2, m = 2.0, np.ones(2, np.float32)
    ori = np.array([[1, 0, 0], [1, 0, 0]], np.float32)
    e = 0.6 * B2
    F = o.rfe(0.0, B2=B2, mask2=m, ori=ori)
    assert F[0] == pytest.approx(-(2 * e / 3) * (t("XG11") - t("XG12")), rel=1e-4)
    assert np.abs(F[1:6]).max() < 1e-9 * abs(F[0])
    # strain transverse to the line of centres (r along z, shear in x-y) exerts no force
    o, _ = _two_spheres(2)
    assert np.abs(o.rfe(1.0)[:3]).max() < 1e-9


def test_rsu_two_sphere_table_columns():
    """S = -R_SU U for two spheres on the x axis with sphere 1 moving: U1 = x gives
    sphere 0 Sxx = -(2/3) XG12 and Syy = +(1/3) XG12; U1 = y gives Sxy = -YG12;
    W1 = z gives Sxy = -YH12; sphere 0's own motion U0 = x gives Sxx = -(2/3) XG11"""
    o, t = _two_spheres(0)
    for k, checks in [(6, {0: -(2 / 3) * t("XG12"), 4: (1 / 3) * t("XG12"), 1: 0.0}),
                      (7, {1: -t("YG12"), 0: 0.0}),
                      (11, {1: -t("YH12"), 0: 0.0}),
                      (0, {0: -(2 / 3) * t("XG11"), 4: (1 / 3) * t("XG11")})]:
        U = np.zeros(12)
        U[k] = 1.0
        S = o.rsu(U)
        for c, want in checks.items():
            assert S[c] == pytest.approx(want, rel=1e-4, abs=1e-9 * max(1.0, abs(t("XG11")))), (k, c)


def test_rse_two_sphere_table_columns():
    """R_SE:E for two spheres: the axisymmetric, mixed and transverse strain modes
    select XM, YM and ZM (Jeffrey 1992's decomposition of M_ijkl): r along x with
    E = e (x x - I/3) gives Sxx = (2e/3)(XM11 + XM12); r along x with 2Exy = 1 gives
    Sxy = (YM11 + YM12)/2; r along z with 2Exy = 1 gives Sxy = (ZM11 + ZM12)/2"""
    o, t = _two_spheres(0)
    S = o.rse(1.0)
    assert S[1] == pytest.approx(0.5 * (t("YM11") + t("YM12")), rel=1e-4)
    B2, m = 2.0, np.ones(2, np.float32)
    ori = np.array([[1, 0, 0], [1, 0, 0]], np.float32)
    e = 0.6 * B2
    S = o.rse(0.0, B2=B2, mask2=m, ori=ori)
    assert S[0] == pytest.approx((2 * e / 3) * (t("XM11") + t("XM12")), rel=1e-4)
    o, t = _two_spheres(2)
    S = o.rse(1.0)
    assert S[1] == pytest.approx(0.5 * (t("ZM11") + t("ZM12")), rel=1e-4)


def test_rfe_rsu_reciprocal():
    """Onsager reciprocity of the grand resistance tensor, R_SU = R_FE^T: for any
    velocities U and per-particle strains E_k, U . (R_FE:E) = sum_k (-R_SU U)_k : E_k
    on a packing (the RFE and RSU code paths share no arithmetic).  Equality holds
    to the rounding of the unit vector, which both kernels form in float as the
    reference does (|r_hat|^2 = 1 to ~1e-7, so S's implied Szz = -Sxx - Syy is
    off by that much); a wrong sign or column is an O(1) error"""
    pos, L = packing(343, 0.3, 17)
    o = _oracle(pos, L)
    rng = np.random.default_rng(5)
    ori = rng.standard_normal((343, 3))
    ori = (ori / np.linalg.norm(ori, axis=1, keepdims=True)).astype(np.float32)
    mask = rng.uniform(0.2, 1.0, 343).astype(np.float32)
    U = rng.standard_normal(6 * 343)
    F = o.rfe(0.7, B2=1.3, mask2=mask, ori=ori)
    S = o.rsu(U)
    Es = _E_tensors(0.7, 1.3, mask, ori)
    lhs = U @ F
    rhs = sum(np.sum(_S5_to_tensor(S[5 * k:5 * k + 5]) * Es[k]) for k in range(343))
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs), (lhs, rhs)


def test_rse_symmetric():
    """R_SE is symmetric: E2 : (R_SE:E1) = E1 : (R_SE:E2) for two fields of
    per-particle strains (shear plus random uniaxial modes), to the float rounding
    of the unit vector as in test_rfe_rsu_reciprocal"""
    pos, L = packing(343, 0.3, 17)
    o = _oracle(pos, L)
    rng = np.random.default_rng(6)
    fields = []
    for shear in (0.3, -0.7):
        ori = rng.standard_normal((343, 3))
        ori = (ori / np.linalg.norm(ori, axis=1, keepdims=True)).astype(np.float32)
        mask = rng.uniform(0.2, 1.0, 343).astype(np.float32)
        fields.append((shear, mask, ori))
    S = [o.rse(sh, B2=0.9, mask2=m, ori=p) for sh, m, p in fields]
    E = [_E_tensors(sh, 0.9, m, p) for sh, m, p in fields]
    a = sum(np.sum(_S5_to_tensor(S[0][5 * k:5 * k + 5]) * E[1][k]) for k in range(343))
    b = sum(np.sum(_S5_to_tensor(S[1][5 * k:5 * k + 5]) * E[0][k]) for k in range(343))
    assert abs(a - b) <= 1e-6 * max(abs(a), 1.0), (a, b)


@pytest.mark.parametrize("gap", [-0.02, -1e-3, 0.0, 0.05, 0.3, 0.49, 0.6])
def test_force_law_closed_form(gap):
    """Stokes_SetForce_manually_kernel (Helper_Stokes.cu:136-160) for a pair on the
    x axis at surface gap g = r - 2 - sqrt(epsq): f = Ham/(12 (g^2 + epsq)) -
    F0 exp(-kappa g) for 0 <= g and r <= 2 + 10/kappa, Ham/(12 epsq) - F0 - k_n |g|
    for g < 0, zero beyond; F0 = 1/ndsr, Ham = F0 beta.  f > 0 pulls sphere 0
    towards sphere 1; the pair force is antisymmetric (Newton's third law) and the
    torque is T_ext along z"""
    ndsr, k_n, kappa, beta, epsq, T_ext = 0.1, 1e4, 20.0, 0.3, 1e-5, 0.25
    rcol = float(np.float32(2.0 + np.sqrt(np.float32(epsq))))
    r = float(np.float32(rcol + gap))
    pos = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]], np.float32)
    o = _oracle(pos, 30.0)
    F = o.force_law(ndsr=ndsr, k_n=k_n, kappa=kappa, beta=beta, epsq=epsq, T_ext=T_ext)
    F0, Ham = 1.0 / ndsr, beta / ndsr
    g = float(np.float32(np.float32(r) - np.float32(rcol)))
    if g < 0:
        f = Ham / (12 * epsq) - F0 - k_n * abs(g)
    elif r <= 2 + 10 / kappa:
        f = Ham / (12 * (g * g + epsq)) - F0 * np.exp(-kappa * g)
    else:
        f = 0.0
    assert F[0] == pytest.approx(f, rel=1e-6, abs=1e-9)
    assert F[6] == pytest.approx(-f, rel=1e-6, abs=1e-9)
    assert np.abs(F[[1, 2, 7, 8]]).max() == 0.0
    assert F[5] == T_ext and F[11] == T_ext and F[3] == 0.0
    if gap == 0.6:
        assert f == 0.0


# ------------------------------------------ Jacobi application of the factor
def test_jacobi_apply_limits():
    """the oracle's k-step Jacobi application of the IC(0) factor: k at or beyond
    the depth of the factor's DAG reproduces the exact triangular solves
    (orc_precond_apply), and k = 0 is the block-diagonal preconditioner
    P^T (L_d L_d^T)^-1 P"""
    pos, L = packing(343, 0.2, 11)
    o = _oracle(pos, L)
    o.build_precond()
    x = np.random.default_rng(3).standard_normal(6 * 343)
    Ld, rows, cols, Lo = o.factor_blocks()
    depth = np.zeros(343, dtype=np.int64)
    for e in np.argsort(rows, kind="stable"):
        depth[rows[e]] = max(depth[rows[e]], depth[cols[e]] + 1)
    exact = o.precond_apply(x)
    got = o.precond_apply_jacobi(x, int(depth.max()) + 1)
    assert np.abs(got - exact).max() <= 1e-10 * np.abs(exact).max()
    y0 = o.precond_apply_jacobi(x, 0)
    D = np.linalg.inv(Ld @ np.transpose(Ld, (0, 2, 1)))
    t = x.reshape(343, 6)[o.perm]
    want = np.empty((343, 6))
    want[o.perm] = np.einsum('pij,pj->pi', D, t)
    assert np.abs(y0 - want.reshape(-1)).max() <= 1e-12 * np.abs(want).max()


# ------------------------------------------------- direct Ewald sum (O2)
def test_direct_ewald_xi_independent_and_hasimoto():
    """the direct Ewald sum (no PSE grid) is independent of the splitting xi to 1e-10
    on a C1 packing (every term converged), and a lone sphere's translational
    self-mobility is Hasimoto's 1 - 2.837297/L + 4.18879/L^3 and its rotational one
    0.75 (1 - (4 pi / 3)/L^3)"""
    from tests._cases import rel_err
    pos, L = packing(343, 0.2, 11)
    o = _oracle(pos, L)
    x = np.random.default_rng(1).standard_normal(11 * 343)
    a, b = o.mobility_direct(x, xi=0.5), o.mobility_direct(x, xi=0.7)
    assert rel_err(a, b) < 1e-10
    L1 = 20.0
    o1 = _oracle(np.zeros((1, 3)), L1, error=1e-6, parity=False)
    e = np.zeros(11)
    e[0] = 1.0
    u = o1.mobility_direct(e)
    assert u[0] == pytest.approx(1.0 - 2.837297 / L1 + 4.18879 / L1 ** 3, rel=1e-6)
    e[0], e[3] = 0.0, 1.0
    assert o1.mobility_direct(e)[3] == pytest.approx(0.75 * (1.0 - 4.18879 / L1 ** 3), rel=2e-6)


def test_pse_converges_to_direct_ewald():
    """the PSE grand mobility (O1, the reference's discretisation) approaches the direct
    Ewald sum as its error tolerance tightens: relative max-norm differences below
    8 x the tolerance and falling with it"""
    from tests._cases import rel_err
    pos, L = packing(343, 0.2, 11)
    x = np.random.default_rng(2).standard_normal(11 * 343)
    ref = _oracle(pos, L).mobility_direct(x)
    errs = [rel_err(_oracle(pos, L, error=tol, parity=False).mobility(x, 3), ref) for tol in (1e-3, 3e-4, 1e-4)]
    assert errs[0] < 8e-3 and errs[1] < 8 * 3e-4 and errs[2] < 8e-4
    assert errs[0] > errs[1] > errs[2]

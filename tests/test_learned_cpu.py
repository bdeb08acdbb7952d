"""CPU checks of the learned factor's restatement (oracle/learned.py) and of the
product-side parameter layout (fsd_amd/learned.py).

The analytic gradient of the oracle's loss is pinned by central finite
differences of its FP64 loss (a directional derivative and single
parameters of every layer); its G by the exact block LDL^T inverse."""
import numpy as np
import pytest

from tests._cases import packing


@pytest.fixture(scope="module")
def dense():
    from oracle.oracle import Oracle
    from oracle.learned import LearnedOracle
    pos, L = packing(128, 0.45, 7)
    orc = Oracle(pos, L)
    orc.build_precond()
    lo = LearnedOracle(orc, k=3)
    rng = np.random.default_rng(0)
    X = rng.standard_normal((2, 17 * 128))
    B = np.array([orc.saddle_multiply(x) for x in X])
    return orc, lo, X, B


def test_layouts_agree():
    from fsd_amd import learned as PL
    from oracle import learned as OL
    assert PL.LAYERS == OL.LAYERS and PL.nparams() == OL.nparams() == 74304


def test_zero_outputs_give_one_step_fill(dense):
    """with zero network outputs the pivots are NPASS passes of D = A_PP + lambda S minus
    0.982 of every lower pair's fill term A F A^T, F = Dt^-1 in the first pass and the
    previous pass's D^-1 in each later one, and W is the pair blocks A minus their one-step
    fill A_PR D_R^-1 A_QR^T over the common lower neighbours R (written out pair by pair)"""
    from oracle.learned import init_weights
    orc, lo, X, B = dense
    c = lo.forward(init_weights(1, scale_out=0.0), emulate=False)
    assert np.abs(c["oute"]).max() == 0 and np.abs(c["outn"]).max() == 0
    W, D, Dinv, failed = lo.assemble(c["oute"], c["outn"])
    assert lo.E > 100 and not failed
    Wref = lo.Ablk.copy()
    low = {}
    for e in range(lo.E):
        low.setdefault(lo.rows[e], {})[lo.cols[e]] = e
    ntri = 0
    for e in range(lo.E):
        P, Q = lo.rows[e], lo.cols[e]
        for R, eq in low.get(Q, {}).items():
            if R in low.get(P, {}):
                Wref[e] -= lo.Ablk[low[P][R]] @ np.linalg.inv(D[R]) @ lo.Ablk[eq].T
                ntri += 1
    assert ntri > 0
    # 1e-5: W's basis takes the table coefficients in FP32, the pair blocks A in FP64
    assert np.abs(W - Wref).max() <= 1e-5 * np.abs(Wref).max()
    # the pivots written out pair by pair from the pattern
    from oracle.learned import gamma
    from oracle.learned import NPASS
    g = gamma(0.0)
    assert abs(g - 0.982) < 1e-4 and NPASS == 4
    prev = lo.Dt.copy()
    for p in range(NPASS):
        Dref = lo.Dt.copy()
        for e in range(lo.E):
            P, Q = lo.rows[e], lo.cols[e]
            Dref[P] -= g * lo.Ablk[e] @ np.linalg.inv(prev[Q]) @ lo.Ablk[e].T
        prev = Dref
    assert all(np.all(lv == 0) for lv in lo.levels_p)
    assert np.abs(D - Dref).max() <= 1e-9 * np.abs(Dref).max()


def test_deep_jacobi_is_exact(dense):
    from oracle.learned import init_weights
    orc, lo, X, B = dense
    c = lo.forward(init_weights(2, scale_out=0.3))
    W, D, Dinv, failed = lo.assemble(c["oute"], c["outn"])
    d = np.random.default_rng(1).standard_normal(6 * lo.N)
    ex = lo.G_exact(W, D, d)
    assert np.abs(lo.apply_G(W, Dinv, d, k=200) - ex).max() < 1e-10 * np.abs(ex).max()
    # a truncated application differs
    assert np.abs(lo.apply_G(W, Dinv, d, k=1) - ex).max() > 1e-6 * np.abs(ex).max()


def test_gradient_finite_differences(dense):
    from oracle.learned import init_weights, layer_offsets
    orc, lo, X, B = dense
    w = init_weights(1, scale_out=0.3).astype(np.float64)
    loss, g = lo.grad(w, X, B, emulate=False)
    rng = np.random.default_rng(5)
    v = rng.standard_normal(w.size)
    eps = 1e-6
    fd = (lo.loss(w + eps * v, X, B, emulate=False) - lo.loss(w - eps * v, X, B, emulate=False)) / (2 * eps)
    assert abs(fd - g @ v) < 1e-6 * abs(fd)
    offs, _ = layer_offsets()
    for wo, bo in offs:           # one weight and one bias of every layer
        for i in (wo + 7, bo + 1):
            e = np.zeros(w.size)
            e[i] = 1e-5
            fdi = (lo.loss(w + e, X, B, emulate=False) - lo.loss(w - e, X, B, emulate=False)) / 2e-5
            assert abs(fdi - g[i]) <= 1e-6 * max(abs(fdi), 1e-3), (i, fdi, g[i])


def test_factor_loss_gradient_finite_differences(dense):
    """the supervised objective (learned factor against IC(0)'s LDL^T blocks): analytic
    gradient against central differences of the FP64 loss, and zero loss when the targets
    are the factor itself"""
    from oracle.learned import init_weights
    orc, lo, X, B = dense
    Dic, Wic = lo.factor_targets(orc)
    w = init_weights(3, scale_out=0.3).astype(np.float64)
    loss, g = lo.factor_loss(w, Dic, Wic, emulate=False, grad=True)
    v = np.random.default_rng(6).standard_normal(w.size)
    # a small step: pivots near singular (fill weights up to 2) have large third derivatives,
    # and the central difference's error falls as eps^2 (2e-5 at 1e-6, 1e-8 at 1e-8)
    eps = 1e-8
    fd = (lo.factor_loss(w + eps * v, Dic, Wic, emulate=False) - lo.factor_loss(w - eps * v, Dic, Wic, emulate=False)) / (2 * eps)
    assert loss > 0 and abs(fd - g @ v) < 1e-6 * abs(fd)
    c = lo.forward(w, emulate=False)
    W, D, Dinv, failed = lo.assemble(c["oute"], c["outn"])
    assert lo.factor_loss(w, D, W, emulate=False) < 1e-20


def _shrunk_weights(seed, a_bias, c_bias):
    """weights whose particle head biases a_bias (pivot shift lambda e^a) and edge fill
    bias c_bias push pivots out of the positive definite range"""
    from oracle.learned import init_weights, layer_offsets, L_HEAD_E, L_HEAD_N
    w = init_weights(seed, scale_out=0.3).astype(np.float64)
    offs, _ = layer_offsets()
    w[offs[L_HEAD_N][1] + 0] = a_bias
    w[offs[L_HEAD_N][1] + 1] = a_bias
    w[offs[L_HEAD_E][1] + 6] = c_bias
    return w


@pytest.fixture(scope="module")
def overfilled(dense):
    """the dense case with every fill term A_PQ Dt_Q^-1 A_PQ^T made 12x larger (Dt^-1
    scaled), so that with a small pivot shift the whole fill makes pivots indefinite (the
    near-field blocks of these packings keep the unscaled pivots positive definite)"""
    import copy
    orc, lo, X, B = dense
    lo4 = copy.deepcopy(lo)
    lo4.Dtinv = 12.0 * lo.Dtinv
    return orc, lo4


def test_pivot_safeguard(overfilled):
    """pivots that a small shift and the whole fill term make indefinite: the assembly
    halves the fill term up to three times and then takes the uncorrected pivot
    A_PP + lambda S; every pivot is positive definite and each is the one its level says"""
    from oracle.learned import gamma, LEVELS, _is_pd
    orc, lo = overfilled
    w = _shrunk_weights(1, -8.0, 8.0)
    c = lo.forward(w, emulate=False)
    W, D, Dinv, failed = lo.assemble(c["oute"], c["outn"])
    print("lambda", lo.lam, "levels per pass", [np.bincount(lv, minlength=LEVELS + 1).tolist() for lv in lo.levels_p])
    assert not failed and np.all(np.bincount(lo.levels_p[0], minlength=LEVELS + 1) > 0)
    S = np.diag([1.0, 1.0, 1.0, 4.0 / 3.0, 4.0 / 3.0, 4.0 / 3.0])
    passes = [(lo.levels_p[p], lo.Dpinv[p - 1] if p else lo.Dtinv, lo.Dp[p]) for p in range(len(lo.Dp))]
    assert np.array_equal(passes[-1][2], D)
    for lv, Finv, Dout in passes:
        F = np.zeros((lo.N, 6, 6))
        for e in range(lo.E):
            F[lo.rows[e]] += gamma(c["oute"][e, 6]) * lo.Ablk[e] @ Finv[lo.cols[e]] @ lo.Ablk[e].T
        for P in range(lo.N):
            assert _is_pd(Dout[P])
            Bp = lo.Adiag[P] + lo.lam * np.diag(np.exp(np.repeat(c["outn"][P, :2], 3))) @ S
            if lv[P] < LEVELS:
                ref = Bp - 0.5 ** lv[P] * F[P]
                assert np.abs(Dout[P] - ref).max() <= 1e-9 * np.abs(ref).max()
                for j in range(lv[P]):      # the levels before it were not positive definite
                    assert not _is_pd(Bp - 0.5 ** j * F[P])
            else:
                assert np.allclose(Dout[P], lo.Dt[P])


def test_safeguard_gradient_finite_differences(overfilled):
    """the supervised objective's gradient where pivots sit at levels 1 .. 4 (scaled
    fill gradient, none at level 4) against central differences"""
    orc, lo = overfilled
    Dic, Wic = lo.factor_targets(orc)
    w = _shrunk_weights(2, -8.0, 8.0)
    loss, g = lo.factor_loss(w, Dic, Wic, emulate=False, grad=True)
    lvs = [lv.copy() for lv in lo.levels_p]
    assert all(lv.max() > 0 for lv in lvs)
    v = np.random.default_rng(8).standard_normal(w.size)
    eps = 1e-7
    fp = lo.factor_loss(w + eps * v, Dic, Wic, emulate=False)
    assert all(np.array_equal(a, b) for a, b in zip(lo.levels_p, lvs))
    fm = lo.factor_loss(w - eps * v, Dic, Wic, emulate=False)
    assert all(np.array_equal(a, b) for a, b in zip(lo.levels_p, lvs))
    fd = (fp - fm) / (2 * eps)
    assert abs(fd - g @ v) < 1e-5 * abs(fd)

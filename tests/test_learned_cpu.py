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
    """with zero network outputs W is the pair blocks A and D = A_PP + lambda S minus
    0.98 of every lower pair's fill term A Dt^-1 A^T"""
    from oracle.learned import init_weights
    orc, lo, X, B = dense
    c = lo.forward(init_weights(1, scale_out=0.0), emulate=False)
    assert np.abs(c["oute"]).max() == 0 and np.abs(c["outn"]).max() == 0
    W, D, Dinv, failed = lo.assemble(c["oute"], c["outn"])
    assert lo.E > 100 and not failed
    assert np.abs(W - lo.Ablk).max() <= 1e-5 * np.abs(lo.Ablk).max()
    # the pivots written out pair by pair from the pattern
    from oracle.learned import gamma
    g = gamma(0.0)
    Dref = lo.Dt.copy()
    for e in range(lo.E):
        P, Q = lo.rows[e], lo.cols[e]
        Dref[P] -= g * lo.Ablk[e] @ np.linalg.inv(lo.Dt[Q]) @ lo.Ablk[e].T
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
    eps = 1e-6
    fd = (lo.factor_loss(w + eps * v, Dic, Wic, emulate=False) - lo.factor_loss(w - eps * v, Dic, Wic, emulate=False)) / (2 * eps)
    assert loss > 0 and abs(fd - g @ v) < 1e-6 * abs(fd)
    c = lo.forward(w, emulate=False)
    W, D, Dinv, failed = lo.assemble(c["oute"], c["outn"])
    assert lo.factor_loss(w, D, W, emulate=False) < 1e-20

"""GPU parity of the learned preconditioner factor against its FP64 restatement
(oracle/learned.py), and its fallbacks to IC(0).

The oracle rounds to BF16 where the GPU does (weights, matrix-product inputs,
pre-activation gradients into input-gradient products), so the remaining
differences are FP32 accumulation order and the rare BF16 rounding flips it
causes; the tolerances below are written for that.  The factor's application
(block Jacobi iterations on W and D^-1) is compared on the GPU's own W and
D^-1, where only FP32 rounding remains (1e-5).
"""
import numpy as np
import pytest

from tests._cases import packing, rel_err, rel_err2, random_vec

pytestmark = pytest.mark.gpu

CASES = {"C1": (343, 0.20, 11), "D1": (512, 0.45, 21)}


def _setup(case, w, k=3, **kw):
    import torch
    from fsd_amd.stokes import Stokes
    from oracle.oracle import Oracle
    from oracle.learned import LearnedOracle
    N, phi, seed = CASES[case]
    pos, L = packing(N, phi, seed)
    print("learned case", case, "N", N, "phi", phi, "seed", seed, flush=True)
    st = Stokes(N, L, precond="learned", jacobi_iters=k, learned_weights=w, **kw)
    p4 = np.zeros((N, 4), dtype=np.float32)
    p4[:, :3] = pos
    dpos = torch.from_numpy(p4.reshape(-1)).cuda()
    st.build_nlist(dpos)
    orc = Oracle(pos, L)
    orc.build_precond()
    return st, orc, LearnedOracle(orc, k=k), dpos, N


def _weights(seed=1, head=0.3):
    from oracle.learned import init_weights
    return init_weights(seed, scale_out=head)


@pytest.mark.parametrize("case", ["C1", "D1"])
def test_learned_factor_parity(case):
    w = _weights()
    st, orc, lo, dpos, N = _setup(case, w)
    info = st.build_precond(dpos)
    assert info["precond_used"] == 1 and info["fallback"] == 0
    g = st.learned_export(lo.E)
    c = lo.forward(w, emulate=True)
    W, D, Dinv, failed = lo.assemble(c["oute"], c["outn"])
    assert not failed
    e_oe = rel_err(g["oute"][:, :7], c["oute"][:, :7]) if lo.E else 0.0
    e_on = rel_err(g["outn"][:, :2], c["outn"][:, :2])
    e_W = rel_err(g["W"], W) if lo.E else 0.0
    e_D = rel_err(g["Dinv"], Dinv)
    print("E", lo.E, "outputs edge/particle", e_oe, e_on, "W", e_W, "Dinv", e_D)
    assert e_oe < 2e-2 and e_on < 2e-2 and e_W < 2e-2 and e_D < 2e-2


@pytest.mark.parametrize("case", ["C1", "D1"])
def test_learned_apply_parity(case):
    """G by k Jacobi iterations each way on the GPU's own factor (1e-5), and, with k at
    least the depth of the pattern, equal to the exact ((D + W) D^-1 (D + W)^T)^-1"""
    import torch
    w = _weights()
    for k in (3, 64):
        st, orc, lo, dpos, N = _setup(case, w, k=k)
        st.build_precond(dpos)
        g = st.learned_export(lo.E)
        x = random_vec(6 * N, 4)
        y = st.precond_apply(torch.from_numpy(x.astype(np.float32)).cuda()).cpu().numpy()
        W, Dinv = g["W"].astype(np.float64), g["Dinv"].astype(np.float64)
        ref = lo.apply_G(W, Dinv, x, k=k)
        e = rel_err(y, ref)
        print("k", k, "apply rel err", e)
        assert e < 1e-5
        if k == 64:
            D = np.linalg.inv(Dinv)
            assert rel_err(ref, lo.G_exact(W, D, x)) < 1e-8


@pytest.mark.parametrize("case", ["C1", "D1"])
def test_learned_gradient_parity(case):
    """loss and gradient over all 74304 parameters against the oracle's reverse sweep;
    the oracle's own gradient is pinned by finite differences (tests/test_learned_cpu.py)"""
    import torch
    w = _weights()
    st, orc, lo, dpos, N = _setup(case, w)
    rng = np.random.default_rng(3)
    X = rng.standard_normal((3, 17 * N)).astype(np.float32)
    B = np.stack([orc.saddle_multiply(x.astype(np.float64)) for x in X]).astype(np.float32)
    dX, dB = torch.from_numpy(X).cuda(), torch.from_numpy(B).cuda()
    loss = st.learned_loss_grad(dpos, dX, dB)
    gg = st.learned_get_weights(grad=True)
    lref, gref = lo.grad(w, X.astype(np.float64), B.astype(np.float64), emulate=True)
    el, eg = abs(loss - lref) / abs(lref), rel_err2(gg, gref)
    print("loss", loss, lref, "rel", el, "grad rel L2", eg, "|g|", np.linalg.norm(gref))
    assert el < 1e-3 and eg < 3e-2


@pytest.mark.parametrize("fallback_jacobi", [0, 8])
def test_learned_pivot_fallback(fallback_jacobi):
    """a non-finite network output (injected at one particle's pivot shift) makes the
    build fall back to IC(0) for that step; the solve is then the IC(0) solve, with exact
    triangular solves or (fallback_jacobi = 8) the same application as precond
    'ichol_jacobi' with 8 iterations (equal to FP32 rounding)"""
    import torch
    from fsd_amd.stokes import Stokes
    w = _weights()
    st, orc, lo, dpos, N = _setup("D1", w, gmres_true_tol=1e-5, gmres_restart=100, fallback_jacobi=fallback_jacobi)
    st.learned_inject_pivot(5)
    info = st.build_precond(dpos)
    assert info["precond_used"] == 0 and info["fallback"] == 1
    assert st.learned_stats()["pivot_fallbacks"] == 1
    b = torch.from_numpy(random_vec(17 * N, 9).astype(np.float32)).cuda()
    x, si = st.saddle_solve(dpos, b)
    assert si["precond_used"] == 0 and si["converged"]
    if fallback_jacobi:
        ref = Stokes(N, st.params.Lx, precond="ichol_jacobi", jacobi_iters=fallback_jacobi)
        ref.build_nlist(dpos)
        ref.build_precond(dpos)
        z = torch.from_numpy(random_vec(6 * N, 10).astype(np.float32)).cuda()
        e = rel_err(st.precond_apply(z).cpu().numpy(), ref.precond_apply(z).cpu().numpy())
        print("IC(0) fallback by Jacobi vs ichol_jacobi", e)
        assert e < 1e-6
    # the next build is learned again
    info = st.build_precond(dpos)
    assert info["precond_used"] == 1 and info["fallback"] == 0


def _fallback_sequence(skip_after, steps=9, calibration=6):
    """steps of the dense packing under fresh random forces (the warm start then
    helps neither factor), the learned factor applied by a single Jacobi iteration
    each way, so that its solves reach IC(0)'s iteration count (exact triangular
    solves at the calibration step) and fall back"""
    import torch
    from fsd_amd.stokes import Stokes, ParticleState
    N, phi, seed = CASES["D1"]
    pos, L = packing(N, phi, seed)
    st = Stokes(N, L, precond="learned", jacobi_iters=1, learned_weights=_weights(head=0.0), gmres_tol=1e-5,
                gmres_restart=100, gmres_maxiter=2000, learned_calibration=calibration,
                learned_skip_after=skip_after)
    state = ParticleState(pos)
    seq = []
    for s in range(steps):
        f = torch.from_numpy(random_vec(6 * N, 100 + s).astype(np.float32) * 1e-2).cuda()
        info = st.integrateStepOne(s, state, force6=f, shear_rate=0.0)
        assert info["converged"]
        seq.append((info["precond_used"], info["fallback"]))
    print("skip_after", skip_after, "(precond_used, fallback) per step", seq, st.learned_stats())
    return seq


def test_learned_skip_after_iteration_fallbacks():
    """after two iteration fallbacks in a row (fallback 2) the steps up to the next
    calibration step use IC(0) without building the learned factor (fallback 3); the
    calibration step (every 6th) resets the count, so the step after it tries the
    learned factor again; with the skip switched off every non-calibration step
    tries it"""
    calibration = 6
    seq = _fallback_sequence(2, calibration=calibration)
    consec = 0
    for s, (used, fb) in enumerate(seq):
        if s % calibration == 0:
            assert (used, fb) == (0, 0)
            consec = 0
            continue
        if consec >= 2:
            assert (used, fb) == (0, 3)
            continue
        assert fb in (0, 2)
        consec = consec + 1 if fb == 2 else 0
    assert sum(fb == 3 for _, fb in seq) > 0, "no step reached the skip: the case does not exercise it"
    assert seq[calibration + 1][1] != 3
    seq0 = _fallback_sequence(0, calibration=calibration)
    assert all(fb != 3 for _, fb in seq0)


def test_learned_brownian_solves_capped():
    """Brownian steps (kT = 1) run three saddle solves: the step's own and the two
    random-finite-difference solves.  With the weak learned application of the test
    above, every one of them is capped at IC(0)'s iteration count for the same solve
    (recorded at the calibration step) and finishes with IC(0) when it reaches it:
    each non-calibration step reports the iteration fallback and converges, and the
    count of iteration fallbacks grows by at least one per such step"""
    import torch
    from fsd_amd.stokes import Stokes, ParticleState
    N, phi, seed = CASES["D1"]
    pos, L = packing(N, phi, seed)
    st = Stokes(N, L, T=1.0, seed=5, precond="learned", jacobi_iters=1, learned_weights=_weights(head=0.0),
                gmres_tol=1e-5, gmres_restart=100, gmres_maxiter=2000, learned_calibration=4, learned_skip_after=0)
    state = ParticleState(pos)
    seq, fbs = [], []
    for s in range(4):
        info = st.integrateStepOne(s, state, force6=None, shear_rate=0.0)
        assert info["converged"]
        seq.append((info["precond_used"], info["fallback"], info["iterations"]))
        fbs.append(st.learned_stats()["iteration_fallbacks"])
    print("(precond_used, fallback, iterations) per step", seq, "iteration fallbacks", fbs)
    assert seq[0][:2] == (0, 0) and fbs[0] == 0
    for s in range(1, 4):
        assert seq[s][:2] == (0, 2)
        assert fbs[s] >= fbs[s - 1] + 1


def test_learned_keep_order_steps():
    """steps that do not calibrate keep the last ordering of the particles while its
    neighbour list stands (no RCM of their own graph): the factor is then built on
    another ordering of the same pattern, the solves converge to the same tolerance,
    and the velocities agree with those of steps that order every graph by its RCM to
    the solve tolerance"""
    import torch
    from fsd_amd.stokes import Stokes, ParticleState
    N, phi, seed = CASES["D1"]
    pos, L = packing(N, phi, seed)
    f = torch.from_numpy(random_vec(6 * N, 77).astype(np.float32) * 1e-2).cuda()
    out = {}
    for keep in (True, False):
        st = Stokes(N, L, precond="learned", jacobi_iters=5, learned_weights=_weights(head=0.0), gmres_tol=1e-6,
                    gmres_true_tol=2e-7, refine=True, gmres_restart=100, learned_calibration=4,
                    learned_keep_order=keep, dt=1e-2)
        state = ParticleState(pos)
        vel, its = [], []
        for s in range(4):
            info = st.integrateStepOne(s, state, force6=f, shear_rate=1.0)
            assert info["converged"] and info["true_residual"] <= 2e-7
            vel.append(state.velocity.cpu().numpy().astype(np.float64))
            its.append(info["iterations"])
        out[keep] = (vel, its, state.positions().astype(np.float64))
        print("keep_order", keep, "iterations", its)
    for s in range(4):
        e = rel_err2(out[True][0][s], out[False][0][s])
        print("step", s, "velocity rel diff", e)
        assert e < 1e-5
    assert np.max(np.abs(out[True][2] - out[False][2])) < 1e-5


def test_trained_factor_gain():
    """the committed trained network against the method without it (zero network
    outputs: the pair blocks with uncorrected pivots, block symmetric Gauss-Seidel)
    and against IC(0): mean GMRES iterations per step to the true residual 1e-4 on a
    held-out N = 4096, phi = 0.45 sheared suspension (training used N = 1024)"""
    import os
    from fsd_amd import learned
    if not os.path.exists(learned.DEFAULT):
        pytest.skip("no weights committed")
    if not learned.load_meta().get("trained", True):
        pytest.skip("the committed weights are the zero-output network (no training run improved on it)")
    res = learned.evaluate([("untrained", learned.zero_head_weights()), ("trained", learned.load_weights())],
                           N=4096, phi=0.45, seed=91, steps=4, warmup=2, log=lambda s: print(s, flush=True))
    print("mean iterations", res)
    assert res["trained"] < res["untrained"]


@pytest.mark.parametrize("case", ["C1", "D1"])
def test_learned_factor_loss_parity(case):
    """the supervised objective (the learned factor's pivots and pair blocks against
    IC(0)'s block LDL^T): loss and gradient against the oracle's (IC(0) in FP64 there,
    FP32 on the GPU: the targets agree to 1e-4).  Heads at 0.03 of He-normal keep every
    pivot at safeguard level 0 in every pass: where a pivot sits near a level boundary
    the gradient jumps with the level, and BF16 rounding of the activations can move
    the outputs across it (test_learned_safeguard_parity covers the other levels)"""
    w = _weights(seed=4, head=0.03)
    st, orc, lo, dpos, N = _setup(case, w)
    loss = st.learned_factor_loss(dpos)
    assert loss is not None, "non-finite network outputs"
    gg = st.learned_get_weights(grad=True)
    Dic, Wic = lo.factor_targets(orc)
    lref, gref = lo.factor_loss(w, Dic, Wic, emulate=True, grad=True)
    el, eg = abs(loss - lref) / abs(lref), rel_err2(gg, gref)
    print("factor loss", loss, lref, "rel", el, "grad rel L2", eg)
    assert el < 1e-3 and eg < 1e-2


@pytest.mark.parametrize("case", ["C1", "D1"])
def test_learned_safeguard_parity(case):
    """pivots at every safeguard level: heads with zero weights and biases that set the
    pivot shift to e^-6 lambda and the fill weight to 2 sigmoid(6) (outputs that do not
    depend on the activations, so BF16 rounding cannot move them).  The GPU's level of
    every pivot in every pass equals the oracle's, and the supervised loss agrees.  At C1
    the supervised and the saddle objective's gradients (and the saddle loss) agree with
    the oracle's through the scaled fill terms of levels 1-3 and the output-free pivots of
    level 4.  At D1 these outputs make the factor nearly singular (saddle loss ~5e6) and
    its gradients ill-conditioned: perturbing the pair blocks by 1e-7 (FP32 rounding)
    moves the FP64 oracle's own gradient by 9 % (no level changes), so only the levels and
    the supervised loss are compared there."""
    import torch
    from oracle.learned import init_weights, layer_offsets, L_HEAD_E, L_HEAD_N, LEVELS
    w = init_weights(1, scale_out=0.0)
    offs, _ = layer_offsets()
    w[offs[L_HEAD_N][1] + 0] = -6.0
    w[offs[L_HEAD_N][1] + 1] = -6.0
    w[offs[L_HEAD_E][1] + 6] = 6.0
    st, orc, lo, dpos, N = _setup(case, w)
    loss = st.learned_factor_loss(dpos)
    assert loss is not None
    gg = st.learned_get_weights(grad=True)
    Dic, Wic = lo.factor_targets(orc)
    lref, gref = lo.factor_loss(w, Dic, Wic, emulate=True, grad=True)
    lv = st.learned_levels()
    counts = [np.bincount(l, minlength=LEVELS + 1).tolist() for l in lo.levels_p]
    flips = [int((lv[p] != lo.levels_p[p]).sum()) for p in range(len(lo.levels_p))]
    print("levels per pass (oracle)", counts, "GPU/oracle differences", flips)
    assert sum(flips) == 0
    assert sum(c[1] + c[2] + c[3] for c in counts) > 0
    el, eg = abs(loss - lref) / abs(lref), rel_err2(gg, gref)
    print("factor loss", loss, lref, "rel", el, "grad rel L2", eg)
    assert el < 1e-3 and (case == "D1" or eg < 1e-2)
    if case == "D1":
        return
    rng = np.random.default_rng(3)
    X = rng.standard_normal((2, 17 * N)).astype(np.float32)
    B = np.stack([orc.saddle_multiply(x.astype(np.float64)) for x in X]).astype(np.float32)
    loss = st.learned_loss_grad(dpos, torch.from_numpy(X).cuda(), torch.from_numpy(B).cuda())
    gg = st.learned_get_weights(grad=True)
    lref, gref = lo.grad(w, X.astype(np.float64), B.astype(np.float64), emulate=True)
    assert all(np.array_equal(st.learned_levels()[p], lo.levels_p[p]) for p in range(len(lo.levels_p)))
    el, eg = abs(loss - lref) / abs(lref), rel_err2(gg, gref)
    print("saddle loss", loss, lref, "rel", el, "grad rel L2", eg)
    assert el < 1e-3 and eg < 3e-2

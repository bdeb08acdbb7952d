"""CPU pins of the oracle's Brownian pieces: Lanczos square roots against
dense eigendecompositions, the fluctuation-dissipation theorem for the
wave-space noise, and the symmetry of the random finite difference."""
import numpy as np
import pytest

from oracle.oracle import Oracle, lanczos_sqrt

POS5 = np.array([[0, 0, 0], [2.05, 0.3, 0], [0.4, 2.1, 0.2], [-1.9, -0.9, 0.5], [3.0, 2.5, -1.0]], float)


def _dense(apply, n):
    A = np.array([apply(np.eye(n)[c]) for c in range(n)]).T
    return 0.5 * (A + A.T)


def _sqrt_apply(A, v):
    lam, Q = np.linalg.eigh(A)
    return Q @ (np.sqrt(np.maximum(lam, 0.0)) * (Q.T @ v))


def test_lanczos_sqrt_real_space_mobility():
    """M_real^(1/2) psi by FP64 Lanczos equals the dense square root; M_real is
    symmetric positive definite"""
    o = Oracle(POS5, 12.0)
    n = 11 * 5
    M = _dense(lambda v: o.mobility(v, parts=1), n)
    assert np.linalg.eigvalsh(M).min() > 0
    psi = np.random.default_rng(3).standard_normal(n)
    y, m = lanczos_sqrt(lambda v: o.mobility(v, parts=1), psi, 1e-12)
    ref = _sqrt_apply(M, psi)
    assert np.linalg.norm(y - ref) <= 1e-7 * np.linalg.norm(ref)


def test_lanczos_sqrt_lubrication():
    """R_FU^(1/2) psi (the near-field Brownian force) equals the dense square root"""
    o = Oracle(POS5, 12.0)
    n = 6 * 5
    R = _dense(o.rfu, n)
    psi = np.random.default_rng(4).standard_normal(n)
    y, m = lanczos_sqrt(o.rfu, psi, 1e-12)
    assert np.linalg.norm(y - _sqrt_apply(R, psi)) <= 1e-9 * np.linalg.norm(y)


def test_wave_noise_fluctuation_dissipation():
    """the wave-space far-field slip driven by unit grid noise has covariance
    2 T / dt M_wave (fluctuation-dissipation), checked by sampling"""
    pos = POS5[:3]
    o = Oracle(pos, 12.0)
    n = 11 * 3
    Mw = _dense(lambda v: o.mobility(v, parts=2), n)
    M = o.pse.Nx * o.pse.Ny * o.pse.Nz
    T, dt, ns = 1.0, 1e-3, 3000
    rng = np.random.default_rng(5)
    S = np.zeros((n, n))
    for _ in range(ns):
        u = o.bro_ff_wave(rng.standard_normal(6 * M).astype(np.float32), T, dt, parity=False)
        S += np.outer(u, u)
    S /= ns
    target = 2 * T / dt * Mw
    d = np.sqrt(np.diag(target))
    dev = (S - target) / np.outer(d, d)
    sd = np.sqrt(2.0 / ns)
    assert abs(np.mean(np.diag(S) / np.diag(target)) - 1.0) < 3 * sd
    assert np.abs(dev).max() < 6 * sd


def test_rfd_even_in_psi():
    """Integrator_RFD's drift is even in psi: flipping psi swaps the two solves
    and their right-hand sides"""
    o = Oracle(POS5, 12.0)
    o.build_precond()
    psi = np.random.default_rng(6).standard_normal(30)
    a = o.rfd_drift(psi, 1.0, 1e-3)
    b = o.rfd_drift(-psi, 1.0, 1e-3)
    assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(a)
    assert np.linalg.norm(a) > 0


def test_adaptive_lanczos_rule_converges_to_sqrt():
    """the sampler's adaptive stopping rule (relative change <= tol) gives the
    square root to about tol on the well-conditioned real-space mobility"""
    from oracle.oracle import lanczos_sqrt_adaptive
    o = Oracle(POS5, 12.0)
    n = 11 * 5
    M = _dense(lambda v: o.mobility(v, parts=1), n)
    psi = np.random.default_rng(7).standard_normal(n)
    y, m = lanczos_sqrt_adaptive(lambda v: o.mobility(v, parts=1), psi, 2, 1e-3)
    ref = _sqrt_apply(M, psi)
    assert np.linalg.norm(y - ref) <= 1e-3 * np.linalg.norm(ref)
    assert 2 <= m < 99

"""CPU pins of the oracle's integration kernels (Integrator.cu:126-438): the
explicit Euler step and the two Runge-Kutta stages against closed forms
(translation, rotation of the director about the vorticity axis, the composed
stage 1 + stage 2 being exact for a constant velocity) and against each other."""
import numpy as np

from oracle.oracle import euler_shear, rk2_stage1, rk2_stage2

L = 20.0


def _state(N, seed):
    rng = np.random.default_rng(seed)
    pos = np.zeros((N, 4), np.float32)
    pos[:, :3] = rng.uniform(-L / 2, L / 2, (N, 3))
    pos[:, 3] = 1.0
    ori = rng.standard_normal((N, 3)).astype(np.float32)
    ori /= np.linalg.norm(ori, axis=1, keepdims=True)
    gb = pos.copy()
    img = np.zeros((N, 3), np.int32)
    return pos, ori, gb, img


def test_stage1_translation_and_half_step_unwrapped():
    """no rotation, no shear: pos_out = pos + u dt (not wrapped, w kept), pos_gb += u dt / 2"""
    pos, ori, gb, _ = _state(16, 1)
    u = np.zeros((16, 6))
    u[:, :3] = np.random.default_rng(2).uniform(-5, 5, (16, 3))
    dt = 0.7
    po, oo, g = rk2_stage1(pos, ori, gb, u, dt, 0.0)
    assert np.allclose(po[:, :3], pos[:, :3] + u[:, :3] * dt, atol=1e-5)
    assert np.array_equal(po[:, 3], pos[:, 3])
    assert np.allclose(g[:, :3], gb[:, :3] + 0.5 * u[:, :3] * dt, atol=1e-5)
    assert np.allclose(oo, ori, atol=1e-6)
    assert np.abs(po[:, :3]).max() > L / 2          # some left the box: stage 1 does not wrap


def test_director_rotation_closed_form():
    """a director in the xy plane turned by omega_z over dt: angle (omega_z - shear/2) dt"""
    N = 8
    pos, ori, gb, img = _state(N, 3)
    th0 = np.linspace(0, 2 * np.pi, N, endpoint=False)
    ori = np.stack([np.cos(th0), np.sin(th0), np.zeros(N)], 1).astype(np.float32)
    u = np.zeros((N, 6))
    u[:, 5] = 0.9
    dt, g = 0.5, 0.4
    th = th0 + (0.9 - g / 2) * dt
    _, oo, _ = rk2_stage1(pos, ori, gb, u, dt, g)
    assert np.allclose(oo, np.stack([np.cos(th), np.sin(th), np.zeros(N)], 1), atol=1e-6)
    _, oe, _, _ = euler_shear(pos, ori, gb, img, u, L, dt, g)
    assert np.allclose(oe, oo, atol=1e-6)


def test_stage2_with_one_state_is_euler():
    """c1 = 1, c2 = 0, c3 = 1: positions (wrapped), images and directors of the explicit
    Euler step; the unwrapped position advances half as far; pos.w is written 0"""
    N = 32
    pos, ori, gb, img = _state(N, 4)
    rng = np.random.default_rng(5)
    u = rng.uniform(-30, 30, (N, 6))
    dt, g = 0.3, 0.8
    pe, oe, ge, ie = euler_shear(pos, ori, gb, img, u, L, dt, g, B1=0.6)
    p2, o2, g2, i2 = rk2_stage2(1.0, pos, ori, 0.0, pos, ori, gb, u, img, L, 1.0, dt, g, B1=0.6)
    assert np.allclose(p2[:, :3], pe[:, :3], atol=2e-5)
    assert np.array_equal(i2, ie) and np.any(ie != 0)
    assert np.allclose(o2, oe, atol=1e-6)
    assert np.all(p2[:, 3] == 0.0)
    assert np.allclose(g2[:, :3] - gb[:, :3], 0.5 * (ge[:, :3] - gb[:, :3]), atol=2e-5)


def test_heun_composition_exact_for_constant_velocity():
    """stage 1, then stage 2 with c1 = c2 = c3 = 1/2 (Heun): x0 + u dt for a constant
    velocity, the director turned by the full angle; pos_gb advanced by u dt in total"""
    N = 16
    pos, ori, gb, img = _state(N, 6)
    pos[:, :3] *= 0.2                                 # stay inside the box after the step
    u = np.zeros((N, 6))
    u[:, :3] = np.random.default_rng(7).uniform(-1, 1, (N, 3))
    dt = 0.25
    p1, o1, g1 = rk2_stage1(pos, ori, gb, u, dt, 0.0)
    p2, o2, g2, i2 = rk2_stage2(0.5, pos, ori, 0.5, p1, o1, g1, u, img, L, 0.5, dt, 0.0)
    assert np.allclose(p2[:, :3], pos[:, :3] + u[:, :3] * dt, atol=1e-5)
    assert np.allclose(g2[:, :3], gb[:, :3] + u[:, :3] * dt, atol=1e-5)
    assert np.allclose(o2, ori, atol=1e-6)
    assert not np.any(i2)

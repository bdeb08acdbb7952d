"""CPU tests of the Python user surface: shear protocols (shear_function.py,
SpecificShearFunction.h) and integrate.PSEv3 argument handling."""
import math

import numpy as np
import pytest

from fsd_amd import shear_function as sf


def _rate_is_strain_derivative(f, dt, steps):
    """central differences of the strain match the shear rate"""
    for t in steps:
        d = (f.get_strain(t + 1) - f.get_strain(t - 1)) / (2 * dt)
        assert d == pytest.approx(f.get_shear_rate(t), rel=2e-3, abs=2e-3 * max(1.0, abs(f.get_shear_rate(t))))


def test_steady():
    f = sf.steady(dt=1e-3, shear_rate=2.5, zero=0)
    assert f.get_shear_rate(17) == 2.5
    assert f.get_strain(1000) == pytest.approx(2.5)
    assert sf.steady(dt=0).get_shear_rate(5) == 0.0


def test_sine():
    dt = 1e-3
    f = sf.sine(dt, shear_rate=1.5, shear_freq=2.0, zero=0)
    assert f.get_shear_rate(0) == pytest.approx(1.5)
    assert f.get_strain(125) == pytest.approx(1.5 / (2 * math.pi * 2.0), rel=1e-6)   # quarter period
    _rate_is_strain_derivative(f, dt, [50, 333, 700])
    with pytest.raises(RuntimeError):
        sf.sine(dt, 0.0, 1.0)
    with pytest.raises(RuntimeError):
        sf.sine(dt, 1.0, -1.0)


def test_chirp():
    dt = 1e-3
    f = sf.chirp(dt, amplitude=0.1, omega_0=1.0, omega_f=20.0, periodT=10.0, zero=0)
    assert f.get_strain(0) == pytest.approx(0.0, abs=1e-12)
    _rate_is_strain_derivative(f, dt, [100, 4000, 9000])
    # instantaneous angular frequency reaches omega_f at periodT (logf precision)
    assert 0.1 * 20.0 >= max(abs(f.get_shear_rate(t)) for t in range(9800, 10000)) * 0.999


def test_tukey_and_windowed():
    dt = 1e-3
    w = sf.tukey_window(dt, periodT=2.0, tukey_param=0.5, zero=0)
    assert w.get_strain(0) == 0.0 and w.get_strain(1000) == 1.0 and w.get_strain(2000) == 0.0
    _rate_is_strain_derivative(w, dt, [100, 300, 1700, 1900])
    base = sf.sine(dt, 1.0, 1.0, zero=0)
    f = sf.windowed(base, w)
    t = 150
    assert f.get_strain(t) == pytest.approx(base.get_strain(t) * w.get_strain(t))
    _rate_is_strain_derivative(f, dt, [150, 1000, 1850])
    with pytest.raises(RuntimeError):
        sf.tukey_window(dt, 1.0, 0.0)
    with pytest.raises(RuntimeError):
        sf.tukey_window(dt, 1.0, 1.5)


def test_psev3_rejects_unknown_nlist_type():
    from fsd_amd.integrate import PSEv3

    class _G:
        N = 10
        system = None
    with pytest.raises(RuntimeError):
        PSEv3(_G(), T=0.0, nlist_type="verlet")


def test_output_files(tmp_path):
    """Stokes::OutputData files: names, one line per particle, the reference's
    column layout, traceless stresslet, overlap.txt appended per output step"""
    from fsd_amd.output import write_output
    rng = np.random.default_rng(0)
    N, L = 8, 10.0
    pos = np.array([[0, 0, 0], [2.05, 0, 0], [0, 2.5, 0], [-3, -3, -3], [4, 4, 1], [1, -4, 2], [-2, 3, -4], [3, -1, 4]],
                   dtype=np.float32)
    ori = np.tile([1.0, 0.0, 0.0], (N, 1))
    vel = rng.standard_normal(11 * N).astype(np.float32)
    nbrs, nn = [], []
    for i in range(N):
        d = pos - pos[i]
        d -= L * np.round(d / L)
        js = [j for j in range(N) if j != i and np.linalg.norm(d[j]) < 5.66]
        nbrs += js
        nn.append(len(js))
    box = (L, L, L, 0.0, 0.0, 0.0)
    for ts in (0, 10):
        write_output(ts, pos, ori, pos, vel, np.array(nn), np.array(nbrs), box, 0.1, 20.0, 1e4, 0.0, 1e-5,
                     directory=str(tmp_path))
    for name, ncol in [("stresslet", 7), ("position", 10), ("velocity", 7), ("interparticle_stresslet", 7)]:
        lines = open(tmp_path / "raw" / ("%s.%010d.txt" % (name, 10))).read().splitlines()
        assert len(lines) == N
        assert all(len(l.split()) == ncol for l in lines)
        assert [int(l.split()[0]) for l in lines] == list(range(N))
    s = np.loadtxt(tmp_path / "raw" / ("stresslet.%010d.txt" % 10))
    assert np.allclose(s[:, 1] + s[:, 4] + s[:, 6], 0.0, atol=1e-3)
    assert np.allclose(s[:, 1], vel[6 * N::5] * N / L ** 3, rtol=2e-3)
    ov = open(tmp_path / "overlap.txt").read().splitlines()
    assert len(ov) == 2 and ov[1].split()[0] == "10"
    # the contacting pair (gap 0.05) gives the minimum gap and a repulsive pair stresslet
    assert float(ov[0].split()[4]) == pytest.approx(0.05, abs=1e-3)
    rep = np.loadtxt(tmp_path / "raw" / ("interparticle_stresslet.%010d.txt" % 0))
    assert rep[0, 1] < 0 and rep[1, 1] < 0


def test_shear_variant_wraps_strain():
    """VariantShearFunction: strain wrapped into [-max, max), 0 before the
    offset, frozen after total_timestep"""
    from fsd_amd.variant import shear_variant
    f = sf.steady(dt=1e-2, shear_rate=1.0, zero=0)
    v = shear_variant(f, total_timestep=300, max_strain=0.5)
    assert v(0) == 0.0
    assert v(30) == pytest.approx(0.3)
    assert v(70) == pytest.approx(-0.3)          # 0.7 wraps to -0.3
    assert v(1000) == v(300) == pytest.approx(f.get_strain(300) - 3.0)
    assert all(-0.5 <= v(t) < 0.5 for t in range(0, 400, 7))

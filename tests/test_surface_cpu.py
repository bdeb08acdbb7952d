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

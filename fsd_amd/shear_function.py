"""Time-dependent shear protocols of the reference's plugin
(shear_function.py:1-120, SpecificShearFunction.h): each object gives the
shear rate and the (unwrapped) strain at a timestep.  The integrator reads
get_shear_rate(timestep) every step, as Stokes::integrateStepOne does
(Stokes.cc:1213).

`zero` is the timestep at which the protocol starts; 'now' takes the current
timestep of the active system (fsd_amd.system.current()).
"""
import math

import numpy as np

_PI = 3.1415926536   # the constant SpecificShearFunction.h uses


def _offset(zero):
    from . import system
    now = system.current_timestep()
    if zero == "now":
        return now
    if zero < 0:
        raise RuntimeError("Error creating shear function: negative zero")
    if zero > now:
        raise RuntimeError("Error creating shear function: zero in the future")
    return int(zero)


def _logf(x):
    """the reference evaluates these logarithms in single precision (logf)"""
    return float(np.log(np.float32(x)))


class _shear_function:
    """shear rate and strain as functions of the timestep"""

    def __init__(self, zero="now"):
        self._offset = _offset(zero)

    def get_shear_rate(self, timestep):
        raise NotImplementedError

    def get_strain(self, timestep):
        raise NotImplementedError

    def get_offset(self):
        return self._offset

    def _elapsed(self, timestep):
        return float(int(timestep) - self._offset)


class steady(_shear_function):
    """constant shear rate (SteadyShearFunction); zero shear by default"""

    def __init__(self, dt, shear_rate=0, zero="now"):
        super().__init__(zero)
        self.dt, self.shear_rate = float(dt), float(shear_rate)

    def get_shear_rate(self, timestep):
        return self.shear_rate

    def get_strain(self, timestep):
        return self.shear_rate * self._elapsed(timestep) * self.dt


class sine(_shear_function):
    """rate = max_rate cos(2 pi f t), strain = max_rate sin(2 pi f t) / (2 pi f)
    (SinShearFunction; f is the real, not angular, frequency)"""

    def __init__(self, dt, shear_rate, shear_freq, zero="now"):
        if shear_rate <= 0:
            raise RuntimeError("Error creating shear function: shear rate must be positive")
        if shear_freq <= 0:
            raise RuntimeError("Error creating shear function: shear frequency must be positive")
        super().__init__(zero)
        self.dt, self.shear_rate, self.freq = float(dt), float(shear_rate), float(shear_freq)

    def _phase(self, timestep):
        return self.freq * 2 * _PI * (self._elapsed(timestep) * self.dt)

    def get_shear_rate(self, timestep):
        return self.shear_rate * math.cos(self._phase(timestep))

    def get_strain(self, timestep):
        return self.shear_rate * math.sin(self._phase(timestep)) / self.freq / 2 / _PI


class chirp(_shear_function):
    """exponential chirp from omega_0 to omega_f (angular) over periodT with
    strain amplitude `amplitude` (ChirpShearFunction)"""

    def __init__(self, dt, amplitude, omega_0, omega_f, periodT, zero="now"):
        super().__init__(zero)
        self.dt, self.amp = float(dt), float(amplitude)
        self.omega_0, self.omega_f, self.periodT = float(omega_0), float(omega_f), float(periodT)

    def _growth(self, timestep):
        return math.exp(self.dt * self._elapsed(timestep) * _logf(self.omega_f / self.omega_0) / self.periodT)

    def get_shear_rate(self, timestep):
        omega = self.omega_0 * self._growth(timestep)
        return self.amp * omega * math.cos(self._phase(timestep))

    def get_strain(self, timestep):
        return self.amp * math.sin(self._phase(timestep))

    def _phase(self, timestep):
        return self.periodT * self.omega_0 / _logf(self.omega_f / self.omega_0) * (self._growth(timestep) - 1)


class tukey_window(_shear_function):
    """Tukey window over [0, periodT] with cosine lobes of relative width
    tukey_param (TukeyWindowFunction); used to window another protocol"""

    def __init__(self, dt, periodT, tukey_param, zero="now"):
        if tukey_param <= 0 or tukey_param > 1:
            raise RuntimeError("Error creating Tukey window function: parameter must be within (0, 1]")
        super().__init__(zero)
        self.dt, self.periodT, self.tp = float(dt), float(periodT), float(tukey_param)
        self.omega = 2 * _PI / self.tp

    def _rel(self, timestep):
        return self._elapsed(timestep) * self.dt / self.periodT

    def get_shear_rate(self, timestep):
        t = self._rel(timestep)
        if t <= 0 or t >= 1 or (self.tp / 2 <= t <= 1 - self.tp / 2):
            return 0.0
        shift = self.tp / 2 if t < 0.5 else 1 - self.tp / 2
        return -math.sin(self.omega * (t - shift)) / 2 * self.omega / self.periodT

    def get_strain(self, timestep):
        t = self._rel(timestep)
        if t <= 0 or t >= 1:
            return 0.0
        if self.tp / 2 <= t <= 1 - self.tp / 2:
            return 1.0
        shift = self.tp / 2 if t < 0.5 else 1 - self.tp / 2
        return (1 + math.cos(self.omega * (t - shift))) / 2


class windowed(_shear_function):
    """strain = base strain x window strain; the rate follows the product rule
    (WindowedFunction)"""

    def __init__(self, function_form, window):
        super().__init__("now")
        self.base, self.window = function_form, window

    def get_shear_rate(self, timestep):
        return (self.base.get_shear_rate(timestep) * self.window.get_strain(timestep)
                + self.base.get_strain(timestep) * self.window.get_shear_rate(timestep))

    def get_strain(self, timestep):
        return self.base.get_strain(timestep) * self.window.get_strain(timestep)

    def get_offset(self):
        return self.base.get_offset()

"""Minimal particle system for driving the integrator without HOOMD: a
periodic box, the particle arrays the FSD step reads and writes (device
resident), a timestep counter and the integration methods to run.  It plays
the part of HOOMD's system definition, particle group and run loop for this
one path; nothing else of HOOMD is provided.
"""
import numpy as np

_current = None


def current():
    return _current


def current_timestep():
    return _current.timestep if _current is not None else 0


class ParticleGroup:
    """all particles of a system (the reference asserts group index == tag)"""

    def __init__(self, system):
        self.system = system
        self.N = system.N


class System:
    def __init__(self, positions, box, dt, orientations=None, timestep=0):
        """positions: N x 3; box: L (cube) or (Lx, Ly, Lz[, xy, xz, yz]);
        dt: timestep length; orientations: N x 3 unit directors (default +x)"""
        global _current
        from .stokes import ParticleState
        self.state = ParticleState(np.asarray(positions, dtype=np.float32), ori=orientations)
        self.N = self.state.N
        self.box = (float(box),) * 3 + (0.0,) * 3 if np.isscalar(box) else tuple(float(v) for v in box) + (0.0,) * (6 - len(box))
        self.dt = float(dt)
        self.timestep = int(timestep)
        self.methods = []
        self.last_info = None
        _current = self

    def group_all(self):
        return ParticleGroup(self)

    def run(self, nsteps):
        """advance nsteps timesteps with every registered integration method"""
        for _ in range(int(nsteps)):
            for m in self.methods:
                self.last_info = m.integrate_step(self.timestep)
            self.timestep += 1
        return self.last_info

    def positions(self):
        return self.state.positions()

    def velocities(self):
        return self.state.velocity.cpu().numpy()

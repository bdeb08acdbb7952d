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
        self.tilt_variant = None
        self.last_info = None
        _current = self

    def group_all(self):
        return ParticleGroup(self)

    def box_resize(self, xy):
        """deform the box each timestep: tilt factor xy = xy(timestep) (a
        variant.shear_variant or any callable); particles keep their Cartesian
        positions (the flow already advects them) and are re-wrapped"""
        self.tilt_variant = xy

    def set_tilt(self, xy):
        """new xy tilt factor for the box, the integrators and the particle images"""
        self.box = (self.box[0], self.box[1], self.box[2], float(xy), self.box[4], self.box[5])
        for m in self.methods:
            m.cpp_method.set_box_tilt(xy=self.box[3], xz=self.box[4], yz=self.box[5])
        self._wrap()

    def _wrap(self):
        """HOOMD BoxDim::wrap of every particle into the current box (images updated)"""
        import torch
        Lx, Ly, Lz, xy, xz, yz = self.box
        p = self.state.pos.view(-1, 4)
        img = self.state.image.view(-1, 3)
        x, y, z = p[:, 0].clone(), p[:, 1].clone(), p[:, 2].clone()
        tilt_x = (xz - xy * yz) * z + xy * y
        hi, lo = x >= 0.5 * Lx + tilt_x, x < -0.5 * Lx + tilt_x
        x = torch.where(hi, x - Lx, torch.where(lo, x + Lx, x))
        img[:, 0] += hi.int() - lo.int()
        tilt_y = yz * z
        hi, lo = y >= 0.5 * Ly + tilt_y, y < -0.5 * Ly + tilt_y
        y = torch.where(hi, y - Ly, torch.where(lo, y + Ly, y))
        x = torch.where(hi, x - Ly * xy, torch.where(lo, x + Ly * xy, x))
        img[:, 1] += hi.int() - lo.int()
        hi, lo = z >= 0.5 * Lz, z < -0.5 * Lz
        z = torch.where(hi, z - Lz, torch.where(lo, z + Lz, z))
        y = torch.where(hi, y - Lz * yz, torch.where(lo, y + Lz * yz, y))
        x = torch.where(hi, x - Lz * xz, torch.where(lo, x + Lz * xz, x))
        img[:, 2] += hi.int() - lo.int()
        p[:, 0], p[:, 1], p[:, 2] = x, y, z

    def run(self, nsteps):
        """advance nsteps timesteps: box deformation first (as HOOMD's updaters),
        then every registered integration method"""
        for _ in range(int(nsteps)):
            if self.tilt_variant is not None:
                xy = float(self.tilt_variant(self.timestep))
                if xy != self.box[3]:
                    self.set_tilt(xy)
            for m in self.methods:
                self.last_info = m.integrate_step(self.timestep)
            self.timestep += 1
        return self.last_info

    def positions(self):
        return self.state.positions()

    def velocities(self):
        return self.state.velocity.cpu().numpy()

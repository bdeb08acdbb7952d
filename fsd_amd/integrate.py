"""The reference's user surface, `integrate.PSEv3` (integrate.py:17-155):
same class name, constructor arguments, defaults, set-up order and
methods, driving the MI355X step through fsd_amd.stokes.Stokes.

HOOMD is not part of this project: `group` is a fsd_amd.system.ParticleGroup
(System.group_all()), and System.run(n) plays HOOMD's run loop.  The
neighbour list is the library's own binned list for every `nlist_type`
(the three HOOMD list types the reference accepts give the same pair set).
Extra keyword arguments not in the reference: gmres_tol, gmres_true_tol,
gmres_restart, gmres_maxiter, parity (see fsd_amd.stokes.Stokes).
"""
import math

from . import shear_function as _sf
from .stokes import Stokes


class PSEv3:
    """Fast Stokesian Dynamics integration method (one overdamped step per
    timestep with hydrodynamic interactions)."""

    def __init__(self, group, T, seed=0, xi=0.5, error=0.001, function_form=None, max_strain=0.5,
                 fileprefix="stresslet", period=0, nlist_type="cell", friction_type="none", h0=0.0,
                 alpha=(0.0, 0.0, 0.0, 0.0, 0.0, 0.0), ndsr=1e-1, kappa=1 / 0.05, k_n=1e4, beta_AF=0., epsq=1e-5,
                 sqm_B1=0., sqm_B2=0., N_mix=1, coef_B1_mask=1.0, coef_B2_mask=1.0, rot_diff=0., T_ext=0.,
                 omega_ext=0., **extra):
        if nlist_type.upper() not in ("CELL", "TREE", "STENCIL"):
            raise RuntimeError("Error constructing neighborlist: valid options are cell, tree, stencil")
        self.system = group.system
        self.group = group
        self.period = int(period)
        self.output_dir = extra.pop("output_dir", ".")
        self._force_params = (ndsr, kappa, k_n, beta_AF, epsq)
        self.rcut = math.sqrt(-math.log(error)) / xi
        self._T = T
        T0 = T(self.system.timestep) if callable(T) else float(T)
        self.cpp_method = Stokes(group.N, self.system.box, T=T0, seed=seed, xi=xi, error=error,
                                 fileprefix=fileprefix, period=self.period, ndsr=ndsr, kappa=kappa, k_n=k_n,
                                 beta_AF=beta_AF, epsq=epsq, sqm_B1=sqm_B1, sqm_B2=sqm_B2, N_mix=N_mix,
                                 coef_B1_mask=coef_B1_mask, coef_B2_mask=coef_B2_mask, rot_diff=rot_diff,
                                 T_ext=T_ext, omega_ext=omega_ext, dt=self.system.dt, max_strain=max_strain,
                                 **extra)
        self.function_form = None
        if function_form is not None:
            self.cpp_method.setShear(function_form, max_strain)
            self.function_form = function_form
        else:
            self.function_form = _sf.steady(dt=0)
            self.cpp_method.setShear(self.function_form, max_strain)
        # set-up order of the reference (integrate.py:131-136)
        self.cpp_method.setParams()
        self.cpp_method.setResistanceTable()
        self.cpp_method.setFriction(friction_type.upper(), h0, alpha)
        self.cpp_method.setSparseMath()
        self.cpp_method.AllocateWorkSpaces()
        self.system.methods.append(self)

    def set_params(self, T=None, function_form=None, max_strain=0.5):
        """change the temperature (float or callable of the timestep) and/or the shear protocol"""
        if T is not None:
            self._T = T
            self.cpp_method.setT(T(self.system.timestep) if callable(T) else float(T))
        if function_form is not None:
            self.function_form = function_form
            self.cpp_method.setShear(function_form, max_strain)

    def stop_shear(self, max_strain=0.5):
        self.function_form = _sf.steady(dt=0)
        self.cpp_method.setShear(self.function_form, max_strain)

    def disable(self):
        if self in self.system.methods:
            self.system.methods.remove(self)

    def integrate_step(self, timestep):
        """one Stokes::integrateStepOne at `timestep`, with the reference's output
        before step 0 and after steps whose timestep + 1 is a multiple of period
        (Stokes.cc:1349-1352, 1669-1672)"""
        if callable(self._T):
            self.cpp_method.setT(self._T(timestep))
        if self.period > 0 and timestep == 0:
            self.cpp_method.build_nlist(self.system.state.pos)
            self._output(0)
        rate = self.function_form.get_shear_rate(timestep)
        info = self.cpp_method.integrateStepOne(timestep, self.system.state, shear_rate=rate)
        self.cpp_method.integrateStepTwo(timestep)
        if self.period > 0 and (timestep + 1) % self.period == 0:
            self._output(timestep + 1)
        return info

    def _output(self, timestep):
        from .output import write_output
        s = self.system.state
        nn, nb = self.cpp_method.get_nlist()
        ndsr, kappa, k_n, beta, epsq = self._force_params
        write_output(timestep, s.pos.view(-1, 4).cpu().numpy(), s.ori.cpu().numpy(),
                     s.pos_gb.view(-1, 4).cpu().numpy(), s.velocity.cpu().numpy(), nn, nb, self.system.box, ndsr,
                     kappa, k_n, beta, epsq, directory=self.output_dir)


# the name some FSD scripts use for the same integrator
stokes = PSEv3

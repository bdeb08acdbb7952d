"""Python mirror of the reference's C++ ``Stokes`` integration method
(Stokes.h:43-101) on top of the C ABI (include/fsd.h).

Device buffers are torch CUDA tensors (torch is used only for device memory);
every call synchronises torch's stream before handing pointers to the
library, and the library drains its own stream before returning.
"""
import ctypes as C

import numpy as np

from . import _lib
from .tables import load_resistance_table


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("fsd_amd needs a GPU: torch.cuda.is_available() is False")
    return torch


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class Stokes:
    """One integrator instance: owns an fsd_ctx (device workspaces, tables,
    warm start, IC(0) relaxer) for N particles in a fixed periodic box.

    Method names follow the reference class: setParams / setResistanceTable /
    setFriction / setSparseMath / AllocateWorkSpaces / integrateStepOne /
    integrateStepTwo / setT / setShear.
    """

    def __init__(self, N, box, T=0.0, seed=0, xi=0.5, error=1e-3, fileprefix="stresslet", period=0,
                 ndsr=1e-1, kappa=1 / 0.05, k_n=1e4, beta_AF=0.0, epsq=1e-5, sqm_B1=0.0, sqm_B2=0.0,
                 N_mix=1, coef_B1_mask=1.0, coef_B2_mask=1.0, rot_diff=0.0, T_ext=0.0, omega_ext=0.0,
                 dt=1e-3, max_strain=0.5, gmres_tol=None, gmres_restart=50, gmres_maxiter=1000,
                 parity=True, gmres_true_tol=0.0, guess="previous", gram_schmidt_passes=2,
                 precond="ichol", jacobi_iters=3, factor_sweeps=0, learned_weights="default",
                 learned_calibration=50, refine=False, fallback_jacobi=0, autotune=False, learned_skip_after=2,
                 learned_keep_order=True):
        self.torch = _torch()
        self.lib = _lib.load()
        L = box if not np.isscalar(box) else (box, box, box, 0.0, 0.0, 0.0)
        L = tuple(L) + (0.0,) * (6 - len(L))
        p = _lib.FsdParams()
        p.N = int(N)
        p.Lx, p.Ly, p.Lz, p.xy, p.xz, p.yz = [float(v) for v in L]
        p.T = float(T); p.seed = int(seed) & 0xFFFFFFFF
        p.xi = float(xi); p.error = float(error); p.max_strain = float(max_strain); p.dt = float(dt)
        p.period = int(period)
        p.ndsr, p.kappa, p.k_n, p.beta_AF, p.epsq = float(ndsr), float(kappa), float(k_n), float(beta_AF), float(epsq)
        p.sqm_B1, p.sqm_B2 = float(sqm_B1), float(sqm_B2)
        p.N_mix = int(N_mix)
        p.coef_B1_mask, p.coef_B2_mask = float(coef_B1_mask), float(coef_B2_mask)
        p.rot_diff, p.T_ext, p.omega_ext = float(rot_diff), float(T_ext), float(omega_ext)
        p.gmres_tol = float(gmres_tol) if gmres_tol is not None else float(error)
        p.gmres_restart = int(gmres_restart); p.gmres_maxiter = int(gmres_maxiter)
        p.rlub = 4.0; p.rp = 2.1; p.r_buff = 0.4
        p.parity = 1 if parity else 0
        p.gmres_true_tol = float(gmres_true_tol)
        guesses = {"previous": 0, "extrapolate": 1, "quadratic": 2}
        if guess not in guesses:
            raise ValueError("guess must be 'previous' (the reference), 'extrapolate' (linear) or 'quadratic'")
        p.guess = guesses[guess]
        if gram_schmidt_passes not in (1, 2):
            raise ValueError("gram_schmidt_passes must be 1 or 2")
        p.gs_passes = gram_schmidt_passes
        modes = {"ichol": 0, "ichol_jacobi": 1, "learned": 2}
        if precond not in modes:
            raise ValueError("precond must be 'ichol' (exact triangular solves, the reference), 'ichol_jacobi' "
                             "or 'learned'")
        p.precond = modes[precond]
        p.jacobi_iters = int(jacobi_iters)
        p.factor_sweeps = int(factor_sweeps)
        if refine and not gmres_true_tol:
            raise ValueError("refine needs gmres_true_tol (the FP64 true residual it refines to)")
        p.refine = 1 if refine else 0
        p.fallback_jacobi = int(fallback_jacobi)
        p.autotune = 1 if autotune else 0
        self.params = p
        self.N = int(N)
        self.fileprefix = fileprefix
        self.shear_function = None
        self.max_strain = max_strain
        self.ctx = C.c_void_p()
        rc = self.lib.fsd_create(C.byref(p), C.byref(self.ctx))
        self._check(rc)
        self.setResistanceTable()
        if p.precond == 2:
            # the learned factor: the committed weights unless the caller passes its own
            # (None: no weights, every step then runs IC(0))
            if learned_weights is not None:
                from . import learned
                w = learned.load_weights() if isinstance(learned_weights, str) and learned_weights == "default" \
                    else np.asarray(learned_weights, dtype=np.float32)
                self.learned_set_weights(w)
            self._check(self.lib.fsd_learned_set_calibration(self.ctx, int(learned_calibration)))
            self._check(self.lib.fsd_learned_set_skip(self.ctx, int(learned_skip_after)))
            self._check(self.lib.fsd_learned_set_keep_order(self.ctx, 1 if learned_keep_order else 0))

    # ---- error handling
    def _check(self, rc):
        if rc < 0:
            msg = self.lib.fsd_last_error(self.ctx).decode() if self.ctx else "create failed"
            raise _lib.FsdError(rc, msg)
        return rc

    def __del__(self):
        try:
            if self.ctx:
                self.lib.fsd_destroy(self.ctx)
                self.ctx = C.c_void_p()
        except Exception:
            pass

    def _sync(self):
        self.torch.cuda.synchronize()

    # ---- reference set-up methods
    def setParams(self):
        """parameters are derived at construction (Stokes.cc:118-696)"""
        return self.pse_info()

    def setResistanceTable(self):
        dist, vals = load_resistance_table()
        self._check(self.lib.fsd_set_resistance_table(self.ctx, dist.ctypes.data_as(C.c_void_p),
                                                      vals.ctypes.data_as(C.c_void_p)))

    def setFriction(self, friction_type="NONE", h0=0.0, alpha=(0.0,) * 6):
        a = np.asarray(alpha, dtype=np.float32)
        self._check(self.lib.fsd_set_friction(self.ctx, friction_type.upper().encode(), float(h0),
                                              a.ctypes.data_as(C.c_void_p)))

    def setSparseMath(self):
        """sparse workspaces are allocated at construction (Stokes_SparseMath.cc:15-163)"""

    def AllocateWorkSpaces(self):
        """workspaces are allocated at construction (Stokes.cc:937-1063)"""

    def setT(self, T):
        self._check(self.lib.fsd_set_temperature(self.ctx, float(T)))

    def setShear(self, shear_func, max_strain=0.5):
        self.shear_function = shear_func
        self.max_strain = max_strain

    def set_box_tilt(self, xy=0.0, xz=0.0, yz=0.0):
        p = self.params
        self._check(self.lib.fsd_set_box(self.ctx, p.Lx, p.Ly, p.Lz, float(xy), float(xz), float(yz)))

    def pse_info(self):
        info = _lib.FsdPseInfo()
        self.lib.fsd_get_pse_info(self.ctx, C.byref(info))
        return {k: getattr(info, k) for k, _ in info._fields_}

    # ---- the step
    def integrateStepOne(self, timestep, state, force6=None, noise_ang=None, shear_rate=None, sync=True):
        """advance `state` (a ParticleState of device tensors) by one step.
        Returns the solve statistics; velocities are left in state.velocity.  sync=False
        skips the device-wide synchronisation before the call (the caller has made the
        inputs ready; used when several contexts step concurrently from host threads)."""
        if shear_rate is None:
            shear_rate = self.shear_function.get_shear_rate(timestep) if self.shear_function else 0.0
        st = _lib.FsdSolveStats()
        if sync:
            self._sync()
        rc = self.lib.fsd_step(self.ctx, int(timestep), float(shear_rate), _ptr(state.pos), _ptr(state.ori),
                               _ptr(state.pos_gb), _ptr(state.image), _ptr(force6), _ptr(noise_ang),
                               _ptr(state.velocity), C.byref(st))
        self._check(rc)
        d = st.as_dict()
        d["converged"] = rc == 0
        return d

    def integrateStepTwo(self, timestep):
        """no second half step (Stokes.cc:1688-1690)"""

    # ---- operator level (tests / benchmarks)
    def build_nlist(self, pos):
        self._sync()
        self._check(self.lib.fsd_build_nlist(self.ctx, _ptr(pos)))

    def get_nlist(self):
        N = self.N
        nn = np.zeros(N, dtype=np.int32)
        tot = C.c_int64(0)
        self._check(self.lib.fsd_get_nlist(self.ctx, nn.ctypes.data_as(C.c_void_p), None, 0, C.byref(tot)))
        nbrs = np.zeros(max(tot.value, 1), dtype=np.int32)
        self._check(self.lib.fsd_get_nlist(self.ctx, nn.ctypes.data_as(C.c_void_p),
                                           nbrs.ctypes.data_as(C.c_void_p), tot.value, C.byref(tot)))
        return nn, nbrs[:tot.value]

    def _out(self, n):
        return self.torch.zeros(n, dtype=self.torch.float32, device="cuda")

    def euler_shear(self, pos4, ori3, pos_gb4, image3, vel6, dt, shear_rate, mask1=None, noise3=None):
        """Integrator_ExplicitEuler_Shear_kernel (Integrator.cu:126-228): advances the
        positions, directors, unwrapped positions and images in place"""
        self._sync()
        self._check(self.lib.fsd_euler_shear(self.ctx, _ptr(pos4), _ptr(ori3), _ptr(pos_gb4), _ptr(image3),
                                             _ptr(vel6), _ptr(mask1), _ptr(noise3), float(dt), float(shear_rate)))

    # ---- second-order Runge-Kutta stages (Integrator.cu:232-438; the reference's step never calls them)
    def rk2_stage1(self, pos4, ori3, pos_gb4, vel6, dt, shear_rate, noise3=None):
        """ExplicitEuler1: returns (pos_out4, ori_out3); pos_gb4 is advanced in place"""
        po = self._out(4 * self.N)
        oo = self._out(3 * self.N)
        self._sync()
        self._check(self.lib.fsd_rk2_stage1(self.ctx, _ptr(pos4), _ptr(ori3), _ptr(po), _ptr(oo), _ptr(pos_gb4),
                                            _ptr(vel6), None, _ptr(noise3), float(dt), float(shear_rate)))
        return po, oo

    def rk2_stage2(self, c1, pos1, ori1, c2, pos2, ori2, pos_gb4, vel6, image3, c3, dt, shear_rate, noise3=None):
        """RK_Shear: returns (pos_out4, ori_out3); pos_gb4 and image3 are updated in place"""
        po = self._out(4 * self.N)
        oo = self._out(3 * self.N)
        self._sync()
        self._check(self.lib.fsd_rk2_stage2(self.ctx, float(c1), _ptr(pos1), _ptr(ori1), float(c2), _ptr(pos2),
                                            _ptr(ori2), _ptr(po), _ptr(oo), _ptr(pos_gb4), _ptr(vel6), None,
                                            _ptr(noise3), _ptr(image3), float(c3), float(dt), float(shear_rate)))
        return po, oo

    def rfu(self, pos, U):
        F = self._out(6 * self.N)
        self._sync()
        self._check(self.lib.fsd_rfu(self.ctx, _ptr(pos), _ptr(U), _ptr(F)))
        return F

    def rfu_noalloc(self, pos, U, F=None):
        """R_FU U through the allocation-free entry point (pair records in the library's
        fixed device array; N <= FSD_NOALLOC_MAX_N)"""
        F = self._out(6 * self.N) if F is None else F
        self._sync()
        self._check(self.lib.fsd_rfu_noalloc(self.ctx, _ptr(pos), _ptr(U), _ptr(F)))
        return F

    def rfe(self, pos, shear_rate, ori=None, F=None):
        F = self._out(6 * self.N) if F is None else F
        self._sync()
        self._check(self.lib.fsd_rfe(self.ctx, _ptr(pos), float(shear_rate), _ptr(ori), _ptr(F)))
        return F

    def rsu(self, pos, U, S=None):
        S = self._out(5 * self.N) if S is None else S
        self._sync()
        self._check(self.lib.fsd_rsu(self.ctx, _ptr(pos), _ptr(U), _ptr(S)))
        return S

    def rse(self, pos, shear_rate, ori=None, S=None):
        S = self._out(5 * self.N) if S is None else S
        self._sync()
        self._check(self.lib.fsd_rse(self.ctx, _ptr(pos), float(shear_rate), _ptr(ori), _ptr(S)))
        return S

    def mobility(self, pos, gF, parts=3):
        gU = self._out(11 * self.N)
        self._sync()
        self._check(self.lib.fsd_mobility(self.ctx, _ptr(pos), _ptr(gF), _ptr(gU), int(parts)))
        return gU

    def build_precond(self, pos):
        st = _lib.FsdSolveStats()
        self._sync()
        self._check(self.lib.fsd_build_precond(self.ctx, _ptr(pos), C.byref(st)))
        return st.as_dict()

    TUNING_KEYS = ("spread", "contract", "xpass", "real_lanes", "real_block", "rfu_lanes", "rfu_block")

    def autotune(self, pos):
        """time the variants of the mobility and R_FU kernels at these positions and keep the
        fastest; returns (choices, best microseconds per apply of the wave part, real part, R_FU)"""
        c = np.zeros(8, dtype=np.int32)
        us = np.zeros(3, dtype=np.float32)
        self._sync()
        self._check(self.lib.fsd_autotune(self.ctx, _ptr(pos), c.ctypes.data_as(C.c_void_p),
                                          us.ctypes.data_as(C.c_void_p)))
        return dict(zip(self.TUNING_KEYS, c.tolist())), dict(zip(("wave", "real", "rfu"), us.tolist()))

    def tuning(self):
        c = np.zeros(8, dtype=np.int32)
        us = np.zeros(3, dtype=np.float32)
        self._check(self.lib.fsd_get_tuning(self.ctx, c.ctypes.data_as(C.c_void_p), us.ctypes.data_as(C.c_void_p)))
        d = dict(zip(self.TUNING_KEYS, c.tolist()))
        d["tuned"] = int(c[7])
        return d, dict(zip(("wave", "real", "rfu"), us.tolist()))

    def set_tuning(self, **choices):
        """force launch shapes: any of TUNING_KEYS (the others keep their current values)"""
        cur = self.tuning()[0]
        cur.update(choices)
        c = np.array([cur[k] for k in self.TUNING_KEYS], dtype=np.int32)
        self._check(self.lib.fsd_set_tuning(self.ctx, c.ctypes.data_as(C.c_void_p)))

    def get_precond(self, with_values=True):
        N = self.N
        nnz = C.c_int64(0)
        pn = np.zeros(N, dtype=np.int32)
        perm = np.zeros(N, dtype=np.int32)
        self._check(self.lib.fsd_get_precond(self.ctx, pn.ctypes.data_as(C.c_void_p), None, 0,
                                             perm.ctypes.data_as(C.c_void_p), None, None, None, None, None, None,
                                             0, C.byref(nnz)))
        plist = np.zeros(int(pn.sum()), dtype=np.int32)
        nz = nnz.value
        rowptr = np.zeros(6 * N + 1, dtype=np.int32)
        col = np.zeros(nz, dtype=np.int32)
        val = np.zeros(nz, dtype=np.float32)
        prow = np.zeros(6 * N + 1, dtype=np.int32)
        pcol = np.zeros(nz, dtype=np.int32)
        Lv = np.zeros(nz, dtype=np.float32)
        vp = lambda a: a.ctypes.data_as(C.c_void_p)
        self._check(self.lib.fsd_get_precond(self.ctx, vp(pn), vp(plist), plist.size, vp(perm), vp(rowptr),
                                             vp(col), vp(val), vp(prow), vp(pcol), vp(Lv), nz, C.byref(nnz)))
        return dict(pn=pn, plist=plist, perm=perm, rowptr=rowptr, col=col, val=val, prow=prow, pcol=pcol, L=Lv)

    def precond_apply(self, x6):
        y = self._out(6 * self.N)
        self._sync()
        self._check(self.lib.fsd_precond_apply(self.ctx, _ptr(x6), _ptr(y)))
        return y

    def saddle_multiply(self, pos, x):
        y = self._out(17 * self.N)
        self._sync()
        self._check(self.lib.fsd_saddle_multiply(self.ctx, _ptr(pos), _ptr(x), _ptr(y)))
        return y

    def saddle_precond(self, b):
        x = self._out(17 * self.N)
        self._sync()
        self._check(self.lib.fsd_saddle_precond(self.ctx, _ptr(b), _ptr(x)))
        return x

    def saddle_solve(self, pos, rhs, x0=None, true_residual=True):
        x = self._out(17 * self.N) if x0 is None else x0.clone()
        st = _lib.FsdSolveStats()
        self._sync()
        rc = self.lib.fsd_saddle_solve(self.ctx, _ptr(pos), _ptr(rhs), _ptr(x), 1 if true_residual else 0,
                                       C.byref(st))
        self._check(rc)
        d = st.as_dict()
        d["converged"] = rc == 0
        return x, d

    # ---- learned preconditioner factor (precond="learned")
    def learned_nparams(self):
        n = C.c_int64(0)
        self._check(self.lib.fsd_learned_nparams(self.ctx, C.byref(n)))
        return n.value

    def learned_set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._check(self.lib.fsd_learned_set_weights(self.ctx, w.ctypes.data_as(C.c_void_p), w.size))

    def learned_get_weights(self, grad=False):
        w = np.zeros(self.learned_nparams(), dtype=np.float32)
        f = self.lib.fsd_learned_get_grad if grad else self.lib.fsd_learned_get_weights
        self._check(f(self.ctx, w.ctypes.data_as(C.c_void_p), w.size))
        return w

    def learned_loss_grad(self, pos, X, B, want_grad=True):
        """loss (and gradient, learned_get_weights(grad=True)) of nsamp samples: X, B are
        [nsamp, 17N] device tensors with B = saddle_multiply(X); None on a non-positive pivot"""
        loss = C.c_double(0.0)
        self._sync()
        rc = self.lib.fsd_learned_loss_grad(self.ctx, _ptr(pos), int(X.shape[0]), _ptr(X), _ptr(B),
                                            1 if want_grad else 0, C.byref(loss))
        self._check(rc)
        return None if rc == _lib.FSD_W_PIVOT else loss.value

    def learned_factor_loss(self, pos, want_grad=True):
        """supervised loss of the learned factor against IC(0)'s at these positions (and its
        gradient, learned_get_weights(grad=True)); None on a non-positive pivot"""
        loss = C.c_double(0.0)
        self._sync()
        rc = self.lib.fsd_learned_factor_loss(self.ctx, _ptr(pos), 1 if want_grad else 0, C.byref(loss))
        self._check(rc)
        return None if rc == _lib.FSD_W_PIVOT else loss.value

    def learned_adam_state(self, m=None, v=None):
        """get (no arguments) or set the Adam moments"""
        n = self.learned_nparams()
        vp = lambda a: a.ctypes.data_as(C.c_void_p)
        if m is None and v is None:
            m, v = np.zeros(n, dtype=np.float32), np.zeros(n, dtype=np.float32)
            self._check(self.lib.fsd_learned_get_adam(self.ctx, vp(m), vp(v), n))
            return m, v
        m = np.ascontiguousarray(m, dtype=np.float32)
        v = np.ascontiguousarray(v, dtype=np.float32)
        self._check(self.lib.fsd_learned_set_adam(self.ctx, vp(m), vp(v), n))

    def learned_revert(self):
        """undo the last Adam step"""
        self._check(self.lib.fsd_learned_revert(self.ctx))

    def learned_adam(self, lr, step):
        self._check(self.lib.fsd_learned_adam(self.ctx, float(lr), int(step)))

    def learned_inject_pivot(self, particle):
        self._check(self.lib.fsd_learned_inject_pivot(self.ctx, int(particle)))

    def learned_stats(self):
        """counters of the learned preconditioner: steps, fallbacks to IC(0) (non-finite
        network outputs, more iterations than IC(0)), calibration steps, and of the last
        factor the pivots built with a reduced fill term and those replaced by the
        uncorrected pivot A_PP + lambda S"""
        out = np.zeros(6, dtype=np.int64)
        self._check(self.lib.fsd_learned_stats(self.ctx, out.ctypes.data_as(C.c_void_p)))
        return dict(steps=int(out[0]), pivot_fallbacks=int(out[1]), iteration_fallbacks=int(out[2]),
                    calibrations=int(out[3]), reduced_fill_pivots=int(out[4]), plain_pivots=int(out[5]))

    def learned_levels(self, npass=4):
        """safeguard level of every pivot in each pass of the last learned factor, [pass, particle]
        in the pattern's order"""
        lv = np.zeros((npass, self.N), dtype=np.int32)
        self._check(self.lib.fsd_learned_levels(self.ctx, lv.ctypes.data_as(C.c_void_p), lv.size))
        return lv

    def learned_export(self, n_lower):
        N = self.N
        oute = np.zeros((n_lower, 32), dtype=np.float32)
        outn = np.zeros((N, 32), dtype=np.float32)
        W = np.zeros((n_lower, 6, 6), dtype=np.float32)
        Dinv = np.zeros((N, 6, 6), dtype=np.float32)
        vp = lambda a: a.ctypes.data_as(C.c_void_p)
        self._check(self.lib.fsd_learned_export(self.ctx, vp(oute), vp(outn), vp(W), vp(Dinv), int(n_lower)))
        return dict(oute=oute, outn=outn, W=W, Dinv=Dinv)

    def force_law(self, pos, T_ext=0.0):
        F = self._out(6 * self.N)
        self._sync()
        self._check(self.lib.fsd_force_law(self.ctx, _ptr(pos), float(T_ext), _ptr(F)))
        return F

    # ---- Brownian motion (T > 0)
    def set_noise(self, ff_real=None, ff_grid=None, nf=None, rfd=None):
        """inject unit-normal noise (device tensors; None = drawn by the library)"""
        self._noise = (ff_real, ff_grid, nf, rfd)      # keep the buffers alive
        self._check(self.lib.fsd_set_noise(self.ctx, _ptr(ff_real), _ptr(ff_grid), _ptr(nf), _ptr(rfd)))

    def brownian_farfield(self, pos, timestep=0):
        out = self._out(11 * self.N)
        self._sync()
        self._check(self.lib.fsd_brownian_farfield(self.ctx, _ptr(pos), int(timestep), _ptr(out)))
        return out

    def brownian_nearfield(self, pos, timestep=0):
        out = self._out(6 * self.N)
        self._sync()
        self._check(self.lib.fsd_brownian_nearfield(self.ctx, _ptr(pos), int(timestep), _ptr(out)))
        return out

    def rfd_drift(self, pos, timestep=0):
        out = self._out(6 * self.N)
        self._sync()
        self._check(self.lib.fsd_rfd_drift(self.ctx, _ptr(pos), int(timestep), _ptr(out)))
        return out

    def get_ewald_table(self):
        info = self.pse_info()
        t = np.zeros(8 * (info["ewald_n"] + 1), dtype=np.float32)
        self._check(self.lib.fsd_get_ewald_table(self.ctx, t.ctypes.data_as(C.c_void_p), t.size))
        return t

    # ---- profiling
    def profile(self, enable=True, classes=None, every=1):
        """time kernel classes with HIP events; classes: names from _lib.PROF_NAMES (None = all);
        every: time one launch in `every` of each class"""
        self._check(self.lib.fsd_profile_every(self.ctx, int(every)))
        if not enable:
            mask = 0
        elif classes is None:
            mask = -1
        else:
            mask = 0
            for c in classes:
                mask |= 1 << _lib.PROF_NAMES.index(c)
        self.lib.fsd_profile_enable(self.ctx, mask)

    def profile_get(self, name):
        n = C.c_int64(0)
        ms = C.c_double(0)
        self._check(self.lib.fsd_profile_get(self.ctx, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def profile_reset(self):
        self.lib.fsd_profile_reset(self.ctx)


class ParticleState:
    """Device-resident particle arrays in the layouts of the reference's
    HOOMD arrays: positions (float4), orientation (stored by the reference in
    the accelerations array), unwrapped positions (in the velocities array),
    images, and the 11N Velocity output."""

    def __init__(self, pos, ori=None, image=None, device="cuda"):
        torch = _torch()
        pos = np.asarray(pos, dtype=np.float32)
        N = pos.shape[0]
        p4 = np.zeros((N, 4), dtype=np.float32)
        p4[:, :3] = pos[:, :3]
        self.N = N
        self.pos = torch.from_numpy(p4.reshape(-1)).to(device)
        self.pos_gb = self.pos.clone()
        if ori is None:
            ori = np.zeros((N, 3), dtype=np.float32)
            ori[:, 0] = 1.0
        self.ori = torch.from_numpy(np.ascontiguousarray(ori, dtype=np.float32).reshape(-1)).to(device)
        img = np.zeros((N, 3), dtype=np.int32) if image is None else np.asarray(image, dtype=np.int32)
        self.image = torch.from_numpy(img.reshape(-1)).to(device)
        self.velocity = torch.zeros(11 * N, dtype=torch.float32, device=device)

    def positions(self):
        return self.pos.view(-1, 4)[:, :3].cpu().numpy()

"""ctypes binding of libfsd.so (declarations in include/fsd.h).

The library is loaded from the package directory; if it is missing the import
of anything that needs it raises, there is no fallback path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libfsd.so")

FSD_OK = 0
FSD_E_PARAM = -1
FSD_E_HIP = -2
FSD_E_CAPACITY = -3
FSD_E_ICHOL = -4
FSD_E_STATE = -5
FSD_W_NOT_CONVERGED = 1
FSD_W_PIVOT = 2

# kernel classes timed by fsd_profile_* (bit k of the enable mask)
PROF_NAMES = ["spread", "green", "contract", "fft", "realspace", "rfu", "trsv", "ichol", "gmres_vec", "nlist",
              "precond_build", "sweep", "plan_host", "finish_host", "ljit", "learned_net", "jacobi"]


class FsdParams(C.Structure):
    _fields_ = [
        ("N", C.c_int32),
        ("Lx", C.c_float), ("Ly", C.c_float), ("Lz", C.c_float),
        ("xy", C.c_float), ("xz", C.c_float), ("yz", C.c_float),
        ("T", C.c_float),
        ("seed", C.c_uint32),
        ("xi", C.c_float),
        ("error", C.c_float),
        ("max_strain", C.c_float),
        ("dt", C.c_float),
        ("period", C.c_int32),
        ("ndsr", C.c_float), ("kappa", C.c_float), ("k_n", C.c_float), ("beta_AF", C.c_float),
        ("epsq", C.c_float),
        ("sqm_B1", C.c_float), ("sqm_B2", C.c_float),
        ("N_mix", C.c_uint32),
        ("coef_B1_mask", C.c_float), ("coef_B2_mask", C.c_float), ("rot_diff", C.c_float),
        ("T_ext", C.c_float), ("omega_ext", C.c_float),
        ("gmres_tol", C.c_float),
        ("gmres_restart", C.c_int32),
        ("gmres_maxiter", C.c_int32),
        ("rlub", C.c_float),
        ("rp", C.c_float),
        ("r_buff", C.c_float),
        ("parity", C.c_int32),
        ("gmres_true_tol", C.c_float),
        ("guess", C.c_int32),
        ("gs_passes", C.c_int32),
        ("precond", C.c_int32),
        ("jacobi_iters", C.c_int32),
        ("factor_sweeps", C.c_int32),
        ("refine", C.c_int32),
        ("fallback_jacobi", C.c_int32),
        ("autotune", C.c_int32),
    ]


class FsdPseInfo(C.Structure):
    _fields_ = [
        ("Nx", C.c_int32), ("Ny", C.c_int32), ("Nz", C.c_int32), ("P", C.c_int32),
        ("gaussm", C.c_float), ("eta", C.c_float), ("rcut", C.c_float),
        ("gridh_x", C.c_float), ("gridh_y", C.c_float), ("gridh_z", C.c_float),
        ("ewald_n", C.c_int32),
        ("self_x", C.c_float), ("self_y", C.c_float),
        ("prefac", C.c_float), ("expfac", C.c_float), ("quadW", C.c_float),
    ]


class FsdSolveStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("restarts", C.c_int32),
        ("residual", C.c_float),
        ("true_residual", C.c_float),
        ("ichol_relaxer", C.c_float),
        ("ichol_attempts", C.c_int32),
        ("nnz", C.c_int64),
        ("nlist_total", C.c_int64),
        ("lub_pairs", C.c_int64),
        ("lower_blocks", C.c_int64),
        ("levels", C.c_int32),
        ("lanczos_ff", C.c_int32),
        ("lanczos_nf", C.c_int32),
        ("sweep_steps", C.c_int32),
        ("initial_residual", C.c_float),
        ("nlist_rebuilt", C.c_int32),
        ("precond_used", C.c_int32),
        ("fallback", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# (name, restype, argtypes) for every symbol fsd.h declares
_P = C.c_void_p
_I = C.c_int32
_F = C.c_float
SIGNATURES = [
    ("fsd_create", _I, [C.POINTER(FsdParams), C.POINTER(_P)]),
    ("fsd_destroy", None, [_P]),
    ("fsd_last_error", C.c_char_p, [_P]),
    ("fsd_get_pse_info", _I, [_P, C.POINTER(FsdPseInfo)]),
    ("fsd_set_resistance_table", _I, [_P, _P, _P]),
    ("fsd_set_friction", _I, [_P, C.c_char_p, _F, _P]),
    ("fsd_get_resistance_table", _I, [_P, _P]),
    ("fsd_set_temperature", _I, [_P, _F]),
    ("fsd_set_box", _I, [_P, _F, _F, _F, _F, _F, _F]),
    ("fsd_get_ewald_table", _I, [_P, _P, _I]),
    ("fsd_step", _I, [_P, C.c_uint32, _F, _P, _P, _P, _P, _P, _P, _P, C.POINTER(FsdSolveStats)]),
    ("fsd_build_nlist", _I, [_P, _P]),
    ("fsd_get_nlist", _I, [_P, _P, _P, C.c_int64, C.POINTER(C.c_int64)]),
    ("fsd_rfu", _I, [_P, _P, _P, _P]),
    ("fsd_rfu_noalloc", _I, [_P, _P, _P, _P]),
    ("fsd_learned_nparams", _I, [_P, C.POINTER(C.c_int64)]),
    ("fsd_learned_set_weights", _I, [_P, _P, C.c_int64]),
    ("fsd_learned_get_weights", _I, [_P, _P, C.c_int64]),
    ("fsd_learned_get_grad", _I, [_P, _P, C.c_int64]),
    ("fsd_learned_get_adam", _I, [_P, _P, _P, C.c_int64]),
    ("fsd_learned_set_adam", _I, [_P, _P, _P, C.c_int64]),
    ("fsd_learned_set_calibration", _I, [_P, _I]),
    ("fsd_learned_set_keep_order", _I, [_P, _I]),
    ("fsd_learned_set_skip", _I, [_P, _I]),
    ("fsd_learned_inject_pivot", _I, [_P, _I]),
    ("fsd_learned_stats", _I, [_P, _P]),
    ("fsd_learned_export", _I, [_P, _P, _P, _P, _P, C.c_int64]),
    ("fsd_learned_levels", _I, [_P, _P, C.c_int64]),
    ("fsd_learned_loss_grad", _I, [_P, _P, _I, _P, _P, _I, C.POINTER(C.c_double)]),
    ("fsd_learned_factor_loss", _I, [_P, _P, _I, C.POINTER(C.c_double)]),
    ("fsd_learned_adam", _I, [_P, _F, _I]),
    ("fsd_learned_revert", _I, [_P]),
    ("fsd_rfe", _I, [_P, _P, _F, _P, _P]),
    ("fsd_rsu", _I, [_P, _P, _P, _P]),
    ("fsd_rse", _I, [_P, _P, _F, _P, _P]),
    ("fsd_mobility", _I, [_P, _P, _P, _P, _I]),
    ("fsd_build_precond", _I, [_P, _P, C.POINTER(FsdSolveStats)]),
    ("fsd_autotune", _I, [_P, _P, _P, _P]),
    ("fsd_get_tuning", _I, [_P, _P, _P]),
    ("fsd_set_tuning", _I, [_P, _P]),
    ("fsd_get_precond", _I, [_P, _P, _P, C.c_int64, _P, _P, _P, _P, _P, _P, _P, C.c_int64,
                             C.POINTER(C.c_int64)]),
    ("fsd_precond_apply", _I, [_P, _P, _P]),
    ("fsd_saddle_multiply", _I, [_P, _P, _P, _P]),
    ("fsd_saddle_precond", _I, [_P, _P, _P]),
    ("fsd_saddle_solve", _I, [_P, _P, _P, _P, _I, C.POINTER(FsdSolveStats)]),
    ("fsd_force_law", _I, [_P, _P, _F, _P]),
    ("fsd_euler_shear", _I, [_P, _P, _P, _P, _P, _P, _P, _P, _F, _F]),
    ("fsd_rk2_stage1", _I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _F, _F]),
    ("fsd_rk2_stage2", _I, [_P, _F, _P, _P, _F, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F, _F, _F]),
    ("fsd_set_noise", _I, [_P, _P, _P, _P, _P]),
    ("fsd_brownian_farfield", _I, [_P, _P, C.c_uint32, _P]),
    ("fsd_brownian_nearfield", _I, [_P, _P, C.c_uint32, _P]),
    ("fsd_rfd_drift", _I, [_P, _P, C.c_uint32, _P]),
    ("fsd_profile_enable", _I, [_P, _I]),
    ("fsd_profile_every", _I, [_P, _I]),
    ("fsd_profile_get", _I, [_P, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    ("fsd_profile_reset", _I, [_P]),
]

_lib = None


def load():
    """Load libfsd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError("libfsd.so is not built; run `python -m fsd_amd.build` first")
    # PyTorch-ROCm brings its own HIP runtime: it is loaded and brings the GPU up before the
    # library (and with it the system HIP runtime) is loaded; in the other order the library
    # finds no device
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    except ImportError:
        pass
    lib = C.CDLL(LIBPATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class FsdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("fsd error %d: %s" % (code, msg))
        self.code = code

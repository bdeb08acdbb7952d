"""Python front end of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the benchmark's
CPU-baseline leg as the checker; the product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None
_P = C.c_void_p


class Box(C.Structure):
    _fields_ = [("Lx", C.c_float), ("Ly", C.c_float), ("Lz", C.c_float),
                ("xy", C.c_float), ("xz", C.c_float), ("yz", C.c_float)]


class Pse(C.Structure):
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int), ("P", C.c_int),
                ("gaussm", C.c_float), ("eta", C.c_float), ("xi", C.c_float), ("error", C.c_float),
                ("rcut", C.c_float), ("gridh", C.c_float * 3), ("ewald_n", C.c_int), ("ewald_dr", C.c_float),
                ("self", C.c_float * 2), ("prefac", C.c_float), ("expfac", C.c_float), ("quadW", C.c_float)]


class Lub(C.Structure):
    _fields_ = [("dist", _P), ("vals", _P), ("table_min", C.c_float), ("table_dr", C.c_float),
                ("rlub", C.c_float), ("rp", C.c_float)]


class Csr(C.Structure):
    _fields_ = [("n", C.c_int), ("nnz", C.c_int), ("rowptr", C.POINTER(C.c_int)),
                ("col", C.POINTER(C.c_int)), ("val", C.POINTER(C.c_double))]


class System(C.Structure):
    _fields_ = [("N", C.c_int), ("pos4", _P), ("box", Box), ("pse", C.POINTER(Pse)), ("ewald_tab", _P),
                ("gridk4", _P), ("nneigh", _P), ("head", _P), ("nbrs", _P), ("lub", C.POINTER(Lub)),
                ("L", C.POINTER(Csr)), ("perm", _P)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        lib.orc_pse_params.argtypes = [C.POINTER(Box), C.c_float, C.c_float, C.c_float, C.POINTER(Pse)]
        lib.orc_gridk.argtypes = [C.POINTER(Box), C.POINTER(Pse), C.c_int, _P]
        lib.orc_nlist.argtypes = [C.c_int, _P, C.POINTER(Box), C.c_float, _P, _P, _P, C.c_int]
        for fn in ("orc_rfu", "orc_rsu"):
            getattr(lib, fn).argtypes = [C.c_int, _P, C.POINTER(Box), _P, _P, _P, C.POINTER(Lub), _P, _P]
        for fn in ("orc_rfe", "orc_rse"):
            getattr(lib, fn).argtypes = [C.c_int, _P, C.POINTER(Box), _P, _P, _P, C.POINTER(Lub), C.c_double,
                                         C.c_double, _P, _P, _P]
        lib.orc_mobility.argtypes = [C.c_int, _P, C.POINTER(Box), C.POINTER(Pse), _P, _P, _P, _P, _P, _P, _P,
                                     C.c_int]
        lib.orc_prune.argtypes = [C.c_int, _P, C.POINTER(Box), _P, _P, _P, C.c_float, _P, _P, _P, _P, C.c_int]
        lib.orc_build_rfu_csr.argtypes = [C.c_int, _P, C.POINTER(Box), _P, _P, _P, _P, C.POINTER(Lub),
                                          C.POINTER(Csr)]
        lib.orc_rcm.argtypes = [C.c_int, _P, _P, _P]
        lib.orc_permute.argtypes = [C.POINTER(Csr), _P, C.c_int, C.POINTER(Csr)]
        lib.orc_ichol.argtypes = [C.POINTER(Csr), C.c_double]
        lib.orc_precond_apply.argtypes = [C.POINTER(Csr), _P, C.c_int, _P, _P]
        lib.orc_csr_free.argtypes = [C.POINTER(Csr)]
        lib.orc_saddle_multiply.argtypes = [C.POINTER(System), _P, _P]
        lib.orc_saddle_precond.argtypes = [C.POINTER(System), _P, _P]
        lib.orc_gmres.argtypes = [C.POINTER(System), _P, _P, C.c_int, C.c_int, C.c_double, C.c_int,
                                  C.POINTER(C.c_double)]
        lib.orc_bro_ff_wave.argtypes = [C.c_int, _P, C.POINTER(Box), C.POINTER(Pse), _P, _P, C.c_int, C.c_double,
                                        C.c_double, _P]
        lib.orc_force_law.argtypes = [C.c_int, _P, C.POINTER(Box), _P, _P, _P, C.c_double, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_double, _P]
        lib.orc_euler_shear.argtypes = [C.c_int, _P, _P, _P, _P, C.POINTER(Box), _P, C.c_double, _P, _P,
                                        C.c_double, C.c_double]
        lib.orc_rk2_stage1.argtypes = [C.c_int, _P, _P, _P, _P, _P, _P, C.c_double, _P, _P, C.c_double, C.c_double]
        lib.orc_rk2_stage2.argtypes = [C.c_int, C.c_double, _P, _P, C.c_double, _P, _P, _P, _P, _P, _P, C.c_double,
                                       _P, _P, _P, C.POINTER(Box), C.c_double, C.c_double, C.c_double]
        lib.fsd_ewald_table_fill.argtypes = [C.c_float, C.c_int, _P]
        lib.fsd_ewald_functions.argtypes = [C.c_double, C.c_double, _P]
        lib.fsd_ewald_self.argtypes = [C.c_float, _P]
        lib.orc_threads.restype = C.c_int
        lib.orc_mobility_direct.argtypes = [C.c_int, _P, C.POINTER(Box), C.c_double, C.c_double, C.c_double,
                                            _P, _P, C.c_int]
        _lib = lib
    return _lib


def threads():
    """OpenMP threads the oracle's parallel loops use"""
    return int(load().orc_threads())


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def make_box(L, xy=0.0, xz=0.0, yz=0.0):
    if np.isscalar(L):
        L = (L, L, L)
    return Box(float(L[0]), float(L[1]), float(L[2]), float(xy), float(xz), float(yz))


class Oracle:
    """Holds one configuration (positions, box, tables, neighbour list) and
    exposes the oracle operators on it with numpy arrays."""

    def __init__(self, pos, L, xi=0.5, error=1e-3, max_strain=0.5, xy=0.0, res_table=None, parity=True,
                 r_buff=0.4, rlub=4.0, rp=2.1):
        self.lib = load()
        pos = np.asarray(pos, dtype=np.float32)
        if pos.shape[1] == 3:
            p4 = np.zeros((pos.shape[0], 4), dtype=np.float32)
            p4[:, :3] = pos
            pos = p4
        self.pos4 = np.ascontiguousarray(pos)
        self.N = self.pos4.shape[0]
        self.box = make_box(L, xy)
        self.pse = Pse()
        rc = self.lib.orc_pse_params(C.byref(self.box), xi, error, max_strain, C.byref(self.pse))
        if rc != 0:
            raise ValueError("orc_pse_params failed: %d" % rc)
        self.nR = self.pse.ewald_n + 1
        self.ewald = np.zeros(8 * self.nR, dtype=np.float32)
        self.lib.fsd_ewald_table_fill(xi, self.nR, _p(self.ewald))
        M = self.pse.Nx * self.pse.Ny * self.pse.Nz
        self.gridk = np.zeros(4 * M, dtype=np.float64)
        self.lib.orc_gridk(C.byref(self.box), C.byref(self.pse), 1 if parity else 0, _p(self.gridk))
        self.rlist = self.pse.rcut + r_buff
        self.build_nlist()
        if res_table is None:
            from fsd_amd.tables import load_resistance_table
            res_table = load_resistance_table()
        self.res_dist = np.ascontiguousarray(res_table[0], dtype=np.float32)
        self.res_vals = np.ascontiguousarray(res_table[1], dtype=np.float32)
        self.lub = Lub(_p(self.res_dist), _p(self.res_vals), 0.0001, 0.004305, rlub, rp)
        self.L = None
        self.perm = None

    # ---- neighbour list
    def build_nlist(self):
        N = self.N
        self.nneigh = np.zeros(N, dtype=np.int32)
        self.head = np.zeros(N, dtype=np.int32)
        cap = max(N * 64, 1024)
        while True:
            self.nbrs = np.zeros(cap, dtype=np.int32)
            tot = self.lib.orc_nlist(N, _p(self.pos4), C.byref(self.box), self.rlist, _p(self.nneigh),
                                     _p(self.head), _p(self.nbrs), cap)
            if tot <= cap:
                break
            cap = tot
        self.nbrs = self.nbrs[:tot].copy()
        return self.nneigh, self.nbrs

    def _nl(self):
        return _p(self.nneigh), _p(self.head), _p(self.nbrs)

    # ---- lubrication
    def rfu(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        F = np.zeros(6 * self.N)
        self.lib.orc_rfu(self.N, _p(self.pos4), C.byref(self.box), *self._nl(), C.byref(self.lub), _p(U), _p(F))
        return F

    def rsu(self, U, S=None):
        U = np.ascontiguousarray(U, dtype=np.float64)
        S = np.zeros(5 * self.N) if S is None else np.array(S, dtype=np.float64)
        self.lib.orc_rsu(self.N, _p(self.pos4), C.byref(self.box), *self._nl(), C.byref(self.lub), _p(U), _p(S))
        return S

    def rfe(self, shear_rate, F=None, B2=0.0, mask2=None, ori=None):
        F = np.zeros(6 * self.N) if F is None else np.array(F, dtype=np.float64)
        m = None if mask2 is None else np.ascontiguousarray(mask2, dtype=np.float32)
        o = None if ori is None else np.ascontiguousarray(ori, dtype=np.float32)
        self.lib.orc_rfe(self.N, _p(self.pos4), C.byref(self.box), *self._nl(), C.byref(self.lub), shear_rate, B2,
                         None if m is None else _p(m), None if o is None else _p(o), _p(F))
        return F

    def rse(self, shear_rate, S=None, B2=0.0, mask2=None, ori=None):
        S = np.zeros(5 * self.N) if S is None else np.array(S, dtype=np.float64)
        m = None if mask2 is None else np.ascontiguousarray(mask2, dtype=np.float32)
        o = None if ori is None else np.ascontiguousarray(ori, dtype=np.float32)
        self.lib.orc_rse(self.N, _p(self.pos4), C.byref(self.box), *self._nl(), C.byref(self.lub), shear_rate, B2,
                         None if m is None else _p(m), None if o is None else _p(o), _p(S))
        return S

    # ---- mobility
    def mobility(self, gF, parts=3):
        gF = np.ascontiguousarray(gF, dtype=np.float64)
        gU = np.zeros(11 * self.N)
        self.lib.orc_mobility(self.N, _p(self.pos4), C.byref(self.box), C.byref(self.pse), _p(self.ewald),
                              _p(self.gridk), *self._nl(), _p(gF), _p(gU), parts)
        return gU

    # ---- preconditioner
    def mobility_direct(self, gF, parts=3, xi=None, tol=1e-14):
        """O2: the FTS grand mobility by the direct Ewald sum (no PSE discretisation),
        converged to ~tol: real space to xi r_cut = sqrt(-ln tol) + 0.5, wave space to
        k_cut = 2 xi sqrt(-ln tol); independent of xi"""
        xi = self.pse.xi if xi is None else xi
        s = np.sqrt(-np.log(tol))
        r_cut, k_cut = (s + 0.5) / xi, 2.0 * xi * s
        gF = np.ascontiguousarray(gF, dtype=np.float64)
        gU = np.zeros(11 * self.N)
        self.lib.orc_mobility_direct(self.N, _p(self.pos4), C.byref(self.box), float(xi), float(r_cut),
                                     float(k_cut), _p(gF), _p(gU), int(parts))
        return gU

    def prune(self):
        N = self.N
        pn = np.zeros(N, dtype=np.int32)
        phead = np.zeros(N + 1, dtype=np.int32)
        pless = np.zeros(N, dtype=np.int32)
        cap = 16 * N
        plist = np.zeros(cap, dtype=np.int32)
        tot = self.lib.orc_prune(N, _p(self.pos4), C.byref(self.box), *self._nl(), self.lub.rp, _p(pn), _p(phead),
                                 _p(plist), _p(pless), cap)
        self.pn, self.phead, self.plist, self.pless = pn, phead, plist[:tot].copy(), pless
        return pn, phead, self.plist, pless

    def build_precond(self, lam=1.0):
        """prune, R_FU^nf CSR, RCM, permute, IC(0) with relaxer doubling"""
        self.prune()
        A = Csr()
        self.lib.orc_build_rfu_csr(self.N, _p(self.pos4), C.byref(self.box), _p(self.pn), _p(self.phead),
                                   _p(self.plist), _p(self.pless), C.byref(self.lub), C.byref(A))
        self.A = A
        self.A_rowptr = np.ctypeslib.as_array(A.rowptr, shape=(A.n + 1,)).copy()
        self.A_col = np.ctypeslib.as_array(A.col, shape=(A.nnz,)).copy()
        self.A_val = np.ctypeslib.as_array(A.val, shape=(A.nnz,)).copy()
        self.perm = np.zeros(self.N, dtype=np.int32)
        self.lib.orc_rcm(self.N, _p(self.phead), _p(self.plist), _p(self.perm))
        attempts = 0
        while True:
            B = Csr()
            self.lib.orc_permute(C.byref(A), _p(self.perm), self.N, C.byref(B))
            if attempts == 0:
                self.B_rowptr = np.ctypeslib.as_array(B.rowptr, shape=(B.n + 1,)).copy()
                self.B_col = np.ctypeslib.as_array(B.col, shape=(B.nnz,)).copy()
                self.B_val = np.ctypeslib.as_array(B.val, shape=(B.nnz,)).copy()
            fail = self.lib.orc_ichol(C.byref(B), lam)
            attempts += 1
            if fail == 0:
                break
            self.lib.orc_csr_free(C.byref(B))
            lam *= 2.0
        self.L = B
        self.L_val = np.ctypeslib.as_array(B.val, shape=(B.nnz,)).copy()
        self.lam = lam
        self.attempts = attempts
        return self

    def precond_apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(6 * self.N)
        self.lib.orc_precond_apply(C.byref(self.L), _p(self.perm), self.N, _p(x), _p(y))
        return y

    def factor_blocks(self):
        """the IC(0) factor on the permuted particle graph as dense 6x6 blocks:
        (Ld [N,6,6] diagonal blocks, rows, cols, Lo [nlow,6,6]) for the lower blocks
        (row P > column Q), ordered by row then column"""
        N = self.N
        rp, col, val = self.B_rowptr, self.B_col, self.L_val
        Ld = np.zeros((N, 6, 6))
        blocks = {}
        for r in range(6 * N):
            P, a = divmod(r, 6)
            for e in range(rp[r], rp[r + 1]):
                Q, c = divmod(int(col[e]), 6)
                if Q == P:
                    Ld[P, a, c] = val[e]
                elif Q < P:
                    blocks.setdefault((P, Q), np.zeros((6, 6)))[a, c] = val[e]
        keys = sorted(blocks)
        rows = np.array([k[0] for k in keys], dtype=np.int64)
        cols = np.array([k[1] for k in keys], dtype=np.int64)
        Lo = np.array([blocks[k] for k in keys]).reshape(-1, 6, 6)
        return Ld, rows, cols, Lo

    def precond_apply_jacobi(self, x, k):
        """G x with each triangular solve replaced by k Jacobi iterations on the
        block factor: with Z_P = L_PP^-1 and F_PQ = Z_P L_PQ, y = c - F y from
        y = c = Z t (t = x permuted), u = y - F^T u from u = y, result P^T Z^T u.
        k >= the depth of the factor's DAG gives the exact solves."""
        N = self.N
        Ld, rows, cols, Lo = self.factor_blocks()
        Z = np.linalg.inv(Ld)
        F = Z[rows] @ Lo
        t = np.asarray(x, dtype=np.float64).reshape(N, 6)[self.perm]
        c = np.einsum('pij,pj->pi', Z, t)
        y = c
        for _ in range(k):
            acc = np.zeros_like(c)
            np.add.at(acc, rows, np.einsum('eij,ej->ei', F, y[cols]))
            y = c - acc
        u = y
        for _ in range(k):
            acc = np.zeros_like(c)
            np.add.at(acc, cols, np.einsum('eji,ej->ei', F, u[rows]))
            u = y - acc
        g = np.einsum('pji,pj->pi', Z, u)
        out = np.empty((N, 6))
        out[self.perm] = g
        return out.reshape(-1)

    # ---- saddle point
    def _system(self):
        s = System()
        s.N = self.N
        s.pos4 = _p(self.pos4)
        s.box = self.box
        s.pse = C.pointer(self.pse)
        s.ewald_tab = _p(self.ewald)
        s.gridk4 = _p(self.gridk)
        s.nneigh, s.head, s.nbrs = self._nl()
        s.lub = C.pointer(self.lub)
        s.L = C.pointer(self.L) if self.L is not None else None
        s.perm = _p(self.perm) if self.perm is not None else None
        self._sys = s
        return s

    def saddle_multiply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(17 * self.N)
        self.lib.orc_saddle_multiply(C.byref(self._system()), _p(x), _p(y))
        return y

    def saddle_precond(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(17 * self.N)
        self.lib.orc_saddle_precond(C.byref(self._system()), _p(b), _p(x))
        return x

    def gmres(self, b, x0=None, restart=50, maxit=1000, tol=1e-10, use_precond=True):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(17 * self.N) if x0 is None else np.array(x0, dtype=np.float64)
        r = C.c_double(0)
        it = self.lib.orc_gmres(C.byref(self._system()), _p(b), _p(x), restart, maxit, tol,
                                1 if use_precond else 0, C.byref(r))
        return x, it, r.value

    # ---- Brownian motion
    def bro_ff_wave(self, noise, T, dt, parity=True):
        """wave-space far-field Brownian slip (11N) from grid noise (6M floats)"""
        noise = np.ascontiguousarray(noise, dtype=np.float32)
        gU = np.zeros(11 * self.N)
        self.lib.orc_bro_ff_wave(self.N, _p(self.pos4), C.byref(self.box), C.byref(self.pse), _p(self.gridk),
                                 _p(noise), 1 if parity else 0, T, dt, _p(gU))
        return gU

    def bro_ff_real(self, psi, T, dt, tol=1e-10, m0=None):
        """real-space far-field Brownian slip sqrt(2T/dt) M_real^(1/2) psi (11N)
        (Brownian_FarField.cu:493-808): converged FP64 Lanczos, or with m0 the
        sampler's adaptive rule from m0 vectors at relative tolerance tol"""
        op = lambda v: self.mobility(v, parts=1)
        y = lanczos_sqrt(op, psi, tol)[0] if m0 is None else lanczos_sqrt_adaptive(op, psi, m0, tol)[0]
        return np.sqrt(2.0 * T / dt) * y

    def bro_nf(self, psi, tol=1e-10, m0=None):
        """near-field Brownian force R_FU^(1/2) psi (6N); psi carries its
        sqrt(2T/dt) (Brownian_NearField.cu:112-464, operator Lubrication_RFU_kernel);
        m0 as in bro_ff_real"""
        if m0 is None:
            return lanczos_sqrt(self.rfu, psi, tol)[0]
        return lanczos_sqrt_adaptive(self.rfu, psi, m0, tol)[0]

    def shifted(self, dpos3):
        """a copy at positions wrapped(pos + dpos3) that keeps this configuration's
        neighbour list and preconditioner (what Integrator_RFD evaluates)"""
        import copy
        o = copy.copy(self)
        p4 = self.pos4.copy()
        for i in range(self.N):
            w = (C.c_float * 3)(*[float(p4[i, a] + dpos3[i, a]) for a in range(3)])
            img = (C.c_int * 3)(0, 0, 0)
            self.lib.orc_wrap(C.byref(self.box), w, img)
            p4[i, :3] = list(w)
        o.pos4 = np.ascontiguousarray(p4)
        o.L = self.L
        o.A = None          # the copy does not own the factor
        o._owner = self
        return o

    def rfd_drift(self, psi6, T, eps, tol=1e-11):
        """Integrator_RFD (Integrator.cu:549-718): (T/eps) [U(x + eps psi/2) - U(x - eps psi/2)]
        for the saddle right-hand side [0; 0; -psi]"""
        N = self.N
        b = np.zeros(17 * N)
        b[11 * N:] = -np.asarray(psi6, dtype=np.float64)
        d3 = np.asarray(psi6, dtype=np.float64).reshape(N, 6)[:, :3]
        up = self.shifted(0.5 * eps * d3)
        xp, _, _ = up.gmres(b, tol=tol)
        um = self.shifted(-0.5 * eps * d3)
        xm, _, _ = um.gmres(b, x0=xp, tol=tol)
        return (T / eps) * (xp[11 * N:] - xm[11 * N:])

    # ---- force / integration
    def force_law(self, ndsr=0.1, k_n=1e4, kappa=20.0, beta=0.0, epsq=1e-5, T_ext=0.0):
        F = np.zeros(6 * self.N)
        self.lib.orc_force_law(self.N, _p(self.pos4), C.byref(self.box), *self._nl(), ndsr, k_n, kappa, beta, epsq,
                               T_ext, _p(F))
        return F

    def rhs(self, force6, shear_rate):
        """saddle right-hand side at T = 0 (Integrator.cu:831-844)"""
        N = self.N
        b = np.zeros(17 * N)
        b[6 * N + 1:11 * N:5] += shear_rate
        b[11 * N:] -= np.asarray(force6, dtype=np.float64)
        b[11 * N:] = self.rfe(shear_rate, F=b[11 * N:])
        return b

    def __del__(self):
        if getattr(self, "_owner", None) is not None:
            return          # a shifted copy: the factor belongs to its parent
        try:
            if self.L is not None:
                self.lib.orc_csr_free(C.byref(self.L))
            if getattr(self, "A", None) is not None:
                self.lib.orc_csr_free(C.byref(self.A))
        except Exception:
            pass


def _f32(a, shape):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(shape))


def euler_shear(pos4, ori3, pos_gb4, image3, vel6, L, dt, shear_rate, B1=0.0, mask1=None, noise3=None, xy=0.0):
    """Integrator_ExplicitEuler_Shear_kernel (Integrator.cu:126-228); returns new
    (pos4, ori3, pos_gb4, image3) (FP64 arithmetic, FP32 state as the reference stores it)"""
    N = len(pos4)
    p, o, g = _f32(pos4, (N, 4)).copy(), _f32(ori3, (N, 3)).copy(), _f32(pos_gb4, (N, 4)).copy()
    im = np.ascontiguousarray(np.asarray(image3, dtype=np.int32).reshape(N, 3)).copy()
    v = np.ascontiguousarray(np.asarray(vel6, dtype=np.float64).reshape(N, 6))
    m = None if mask1 is None else _f32(mask1, (N,))
    nz = None if noise3 is None else _f32(noise3, (N, 3))
    box = make_box(L, xy)
    load().orc_euler_shear(N, _p(p), _p(o), _p(g), _p(im), C.byref(box), _p(v), B1,
                           None if m is None else _p(m), None if nz is None else _p(nz), dt, shear_rate)
    return p, o, g, im


def rk2_stage1(pos4, ori3, pos_gb4, vel6, dt, shear_rate, B1=0.0, mask1=None, noise3=None):
    """Integrator_ExplicitEuler1_Shear_kernel (Integrator.cu:232-326); returns
    (pos_out4, ori_out3, pos_gb4)"""
    N = len(pos4)
    p, o = _f32(pos4, (N, 4)), _f32(ori3, (N, 3))
    g = _f32(pos_gb4, (N, 4)).copy()
    po, oo = np.zeros((N, 4), np.float32), np.zeros((N, 3), np.float32)
    v = np.ascontiguousarray(np.asarray(vel6, dtype=np.float64).reshape(N, 6))
    m = None if mask1 is None else _f32(mask1, (N,))
    nz = None if noise3 is None else _f32(noise3, (N, 3))
    load().orc_rk2_stage1(N, _p(p), _p(o), _p(po), _p(oo), _p(g), _p(v), B1, None if m is None else _p(m),
                          None if nz is None else _p(nz), dt, shear_rate)
    return po, oo, g


def rk2_stage2(c1, pos1, ori1, c2, pos2, ori2, pos_gb4, vel6, image3, L, c3, dt, shear_rate, B1=0.0, mask1=None,
               noise3=None, xy=0.0):
    """Integrator_RK_Shear_kernel (Integrator.cu:329-438); returns (pos_out4, ori_out3, pos_gb4, image3)"""
    N = len(pos1)
    a, oa, b, ob = _f32(pos1, (N, 4)), _f32(ori1, (N, 3)), _f32(pos2, (N, 4)), _f32(ori2, (N, 3))
    g = _f32(pos_gb4, (N, 4)).copy()
    im = np.ascontiguousarray(np.asarray(image3, dtype=np.int32).reshape(N, 3)).copy()
    po, oo = np.zeros((N, 4), np.float32), np.zeros((N, 3), np.float32)
    v = np.ascontiguousarray(np.asarray(vel6, dtype=np.float64).reshape(N, 6))
    m = None if mask1 is None else _f32(mask1, (N,))
    nz = None if noise3 is None else _f32(noise3, (N, 3))
    box = make_box(L, xy)
    load().orc_rk2_stage2(N, c1, _p(a), _p(oa), c2, _p(b), _p(ob), _p(po), _p(oo), _p(g), _p(v), B1,
                          None if m is None else _p(m), None if nz is None else _p(nz), _p(im), C.byref(box), c3, dt,
                          shear_rate)
    return po, oo, g, im


def friction_type1(dist, vals, h0, alpha, parity=True):
    """Stokes::setFriction TYPE1 (Stokes.cc:1132-1191) on a copy of the table
    values: for gaps h <= h0 add alpha_k (2 h^2/h0^3 - 3 h/h0^2 + 1/h) with
    signs (+, -, -, +, +, +) to columns 3, 4, 5, 6, 9, 10 (the columns the
    reference's code writes; parity=False: the named YA11, YA12, YB11, YB12,
    YC11, YC12 columns 2, 3, 4, 5, 8, 9)"""
    v = np.array(vals, dtype=np.float32).reshape(1000, 22).copy()
    cols = [3, 4, 5, 6, 9, 10] if parity else [2, 3, 4, 5, 8, 9]
    sgn = [1.0, -1.0, -1.0, 1.0, 1.0, 1.0]
    h0 = np.float32(h0)
    for ii in range(1000):
        h = np.float32(float(dist[ii]) - 2.0)
        if not h <= h0:
            continue
        h2, h02 = np.float32(h * h), np.float32(h0 * h0)
        h03 = np.float32(h0 * h02)
        poly = 2.0 / float(h03) * float(h2) - 3.0 / float(h02) * float(h) + 1.0 / float(h)
        for k in range(6):
            v[ii, cols[k]] = np.float32(v[ii, cols[k]] + np.float32(sgn[k]) * np.float32(float(np.float32(alpha[k])) * poly))
    return v.reshape(-1)


def lanczos_sqrt(apply, psi, tol=1e-10, mmax=400):
    """|psi| V_m T_m^(1/2) e_1 by Lanczos with full FP64 reorthogonalisation
    (Chow & Saad 2014), iterated until the relative change of the result is
    below tol.  Returns (y, m)."""
    psi = np.asarray(psi, dtype=np.float64)
    nrm = np.linalg.norm(psi)
    if nrm == 0:
        return np.zeros_like(psi), 0
    V = [psi / nrm]
    alpha, beta = [], []
    y_old = None
    for m in range(1, mmax + 1):
        w = apply(V[-1])
        a = float(V[-1] @ w)
        w = w - a * V[-1] - (beta[-1] * V[-2] if beta else 0.0)
        for v in V:
            w -= (v @ w) * v
        alpha.append(a)
        T = np.diag(alpha) + np.diag(beta, 1) + np.diag(beta, -1)
        lam, Q = np.linalg.eigh(T)
        y = nrm * (np.array(V).T @ (Q @ (np.sqrt(np.maximum(lam, 0.0)) * Q[0])))
        b = np.linalg.norm(w)
        if y_old is not None and np.linalg.norm(y - y_old) <= tol * np.linalg.norm(y):
            return y, m
        if b < 1e-14 * abs(a):
            return y, m
        y_old = y
        beta.append(b)
        V.append(w / b)
    return y, mmax


def lanczos_sqrt_adaptive(apply, psi, m0, tol, mmax=100):
    """the Chow-Saad sampler's stopping rule (Brownian_FarField.cu:597-808):
    m0 Lanczos vectors, then one more at a time until the relative change of
    |psi| V_m T_m^(1/2) e_1 is at most tol or the basis holds mmax - 1
    vectors (FP64, full reorthogonalisation).  Returns (y, m - 1)."""
    psi = np.asarray(psi, dtype=np.float64)
    nrm = np.linalg.norm(psi)
    V = [psi / nrm]
    alpha, beta = [], []

    def step():
        w = apply(V[-1])
        a = float(V[-1] @ w)
        for _ in range(2):
            for v in V:
                w -= (v @ w) * v
        alpha.append(a)
        b = np.linalg.norm(w)
        if b < 1e-8:
            return False
        beta.append(b)
        V.append(w / b)
        return True

    def combine(m):
        T = np.diag(alpha[:m]) + np.diag(beta[:m - 1], 1) + np.diag(beta[:m - 1], -1)
        lam, Q = np.linalg.eigh(T)
        return nrm * (np.array(V[:m]).T @ (Q @ (np.sqrt(np.maximum(lam, 0.0)) * Q[0])))

    m = max(1, min(m0, mmax - 2))
    ok = True
    for j in range(m):
        ok = step()
        if not ok:
            m = j + 1
            break
    y = combine(m)
    err = 1.0
    while ok and err > tol and m < mmax - 1:
        ok = step()
        y1 = combine(m + 1)
        err = np.linalg.norm(y1 - y) / np.linalg.norm(y1)
        y = y1
        m += 1
    return y, m - 1


def pse_params(L, xi=0.5, error=1e-3, max_strain=0.5):
    lib = load()
    box = make_box(L)
    p = Pse()
    rc = lib.orc_pse_params(C.byref(box), xi, error, max_strain, C.byref(p))
    if rc != 0:
        raise ValueError("orc_pse_params failed: %d" % rc)
    return p


def ewald_functions(r, xi):
    lib = load()
    out = np.zeros(7)
    lib.fsd_ewald_functions(float(r), float(xi), _p(out))
    return out


def ewald_table(xi, nR):
    lib = load()
    t = np.zeros(8 * nR, dtype=np.float32)
    lib.fsd_ewald_table_fill(xi, nR, _p(t))
    return t

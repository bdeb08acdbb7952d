"""CPU restatement of the learned preconditioner factor (test infrastructure only).

FP64 numpy version of what fsd_amd/csrc/k_learned.hip computes: the edge and
particle features, the 3-layer message-passing network (hidden 64), the block
LDL^T factor it defines on IC(0)'s pattern (pivots from NPASS parallel passes of
IC(0)'s pivot recurrence with the positive-definite safeguard, pair blocks with
their one-step fill correction), the preconditioner G applied by k block Jacobi
iterations each way, the training losses and their gradients with respect to
every network parameter.

The learned factor has no counterpart in the reference; what it replaces is
the IC(0) factorisation of the saddle preconditioner (Precondition.cu:1116-1307,
1448-1489) on the pattern that Precondition.cu:530-836 builds, and the
preconditioner it feeds is Saddle_Preconditioner (Saddle.cu:171-264).

`emulate=True` rounds to BF16 (round to nearest even) exactly where the GPU
does: the weights, every matrix-product input, and the pre-activation
gradients that enter the input-gradient products.  `emulate=False` is plain
FP64; the finite-difference tests of the analytic gradient use it.

Parity of this module's gradient is pinned by central finite differences of
its own FP64 losses, also where pivots sit at every safeguard level
(tests/test_learned_cpu.py); its factor by the zero-output case written out
pair by pair (the pivot passes and the fill correction); its G by the identity
G = ((D + W) D^-1 (D + W)^T)^-1 for k at least the depth of the pattern.
"""
import numpy as np

H, FIN, OE, ON = 64, 32, 32, 32
LAYERS = [(FIN, H, 1), (FIN, H, 1),
          (3 * H, H, 1), (2 * H, H, 1), (3 * H, H, 1), (2 * H, H, 1), (3 * H, H, 1), (2 * H, H, 1),
          (3 * H, OE, 0), (H, ON, 0)]
L_ENC_N, L_ENC_E, L_MSG, L_UPD, L_HEAD_E, L_HEAD_N = 0, 1, 2, 3, 8, 9
S6 = np.array([1.0, 1.0, 1.0, 4.0 / 3.0, 4.0 / 3.0, 4.0 / 3.0])


def layer_offsets():
    offs, o = [], 0
    for ki, no, _ in LAYERS:
        offs.append((o, o + ki * no))
        o += ki * no + no
    return offs, o


def nparams():
    return layer_offsets()[1]


def unpack(w):
    offs, _ = layer_offsets()
    out = []
    for (wo, bo), (ki, no, _) in zip(offs, LAYERS):
        out.append((np.asarray(w[wo:wo + ki * no], dtype=np.float64).reshape(ki, no),
                    np.asarray(w[bo:bo + no], dtype=np.float64)))
    return out


def bf16(x):
    """round float32 values to bfloat16 (nearest even), returned as float64"""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


GAMMA0 = -0.036004   # logit(0.491): gamma(0) = 0.982
LEVELS = 4   # pivot safeguard: fill scales 1, 1/2, 1/4, 1/8, then the uncorrected pivot
NPASS = 4    # parallel passes of the pivot recurrence


def _is_pd(M):
    try:
        np.linalg.cholesky(M)
        return True
    except np.linalg.LinAlgError:
        return False


def gamma(c):
    """fill weight of a pair from its raw output: 2 sigmoid(c + c0), in (0, 2)"""
    return 2.0 / (1.0 + np.exp(-(c + GAMMA0)))


def gamma_d(c):
    sg = 0.5 * gamma(c)
    return 2.0 * sg * (1.0 - sg)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_d(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def eps_r(r):
    """eps.r as the pair blocks use it: E[a][c] = eps_{a c k} r_k (k_precond.hip blocks_kernel)"""
    E = np.zeros(r.shape[:-1] + (3, 3))
    E[..., 0, 1], E[..., 0, 2] = r[..., 2], -r[..., 1]
    E[..., 1, 0], E[..., 1, 2] = -r[..., 2], r[..., 0]
    E[..., 2, 0], E[..., 2, 1] = r[..., 1], -r[..., 0]
    return E


def basis(geo):
    """the six basis blocks [E, 6, 6, 6] of every pair: coefficient times r r, I - r r or
    eps.r in the UU, UW, WU and WW sub-blocks"""
    r = geo[:, 0:3]
    ph = geo[:, 3:8]
    rr = np.einsum('ea,ec->eac', r, r)
    im = np.eye(3)[None] - rr
    E3 = eps_r(r)
    Bk = np.zeros((geo.shape[0], 6, 6, 6))
    Bk[:, 0, 0:3, 0:3] = ph[:, 0, None, None] * rr
    Bk[:, 1, 0:3, 0:3] = ph[:, 1, None, None] * im
    Bk[:, 2, 0:3, 3:6] = ph[:, 2, None, None] * E3
    Bk[:, 3, 3:6, 0:3] = ph[:, 2, None, None] * E3
    Bk[:, 4, 3:6, 3:6] = ph[:, 3, None, None] * rr
    Bk[:, 5, 3:6, 3:6] = ph[:, 4, None, None] * im
    return Bk


def table_lookup(dist, dist_tab, vals, table_min=1e-4, table_dr=0.004305):
    """the ten RFU columns (XA11 .. YC12) at each distance with the preconditioner's
    contact clamp (table row 0 at r <= 2 + h_min, Precondition.cu:612-624) and the
    lerp of Lubrication.cu:172-231"""
    dist = np.asarray(dist, dtype=np.float32)
    V = vals.reshape(1000, 22)
    out = np.zeros((dist.shape[0], 10))
    for n, d in enumerate(dist):
        if float(d) <= 2.0 + float(np.float32(table_min)):
            out[n] = V[0, :10]
            continue
        arg = np.float32((float(d) - 2.0) / float(np.float32(table_min)))
        ind = int(np.float32(np.log10(arg)) / np.float32(table_dr))
        ind = max(ind, 0)
        if ind >= 999:
            out[n] = V[999, :10]
            continue
        dl, du = dist_tab[ind], dist_tab[ind + 1]
        fac = (np.float32(d) - dl) / (du - dl)
        out[n] = V[ind, :10] + (V[ind + 1, :10] - V[ind, :10]) * fac
    return out


class LearnedOracle:
    """The learned factor on the pattern of an oracle whose build_precond has run."""

    def __init__(self, orc, lam=None, k=3):
        self.N, self.k = orc.N, k
        N = self.N
        self.perm = orc.perm.astype(np.int64)
        rp, col, val = orc.B_rowptr, orc.B_col, orc.B_val
        Ad = np.zeros((N, 6, 6))
        blocks = {}
        for r in range(6 * N):
            P, a = divmod(r, 6)
            for e in range(rp[r], rp[r + 1]):
                Q, c = divmod(int(col[e]), 6)
                if Q == P:
                    Ad[P, a, c] = val[e]
                elif Q < P:
                    blocks.setdefault((P, Q), np.zeros((6, 6)))[a, c] = val[e]
        keys = sorted(blocks)
        self.rows = np.array([q[0] for q in keys], dtype=np.int64)
        self.cols = np.array([q[1] for q in keys], dtype=np.int64)
        self.Ablk = np.array([blocks[q] for q in keys]).reshape(-1, 6, 6)
        self.Adiag = Ad
        self.E = len(keys)
        # geometry of every lower pair (row particle i, column particle j), float32 like the GPU
        pos = orc.pos4[:, :3].astype(np.float32)
        Lbox = np.float32(orc.box.Lx)
        i, j = self.perm[self.rows], self.perm[self.cols]
        d = (pos[j] - pos[i]).astype(np.float32)
        d = (d - Lbox * np.rint(d / Lbox).astype(np.float32)).astype(np.float32)
        dist = np.sqrt((d.astype(np.float32) ** 2).sum(axis=1, dtype=np.float32)).astype(np.float32)
        v = table_lookup(dist, orc.res_dist, orc.res_vals) if self.E else np.zeros((0, 10))
        self.geo = np.zeros((self.E, 8))
        if self.E:
            self.geo[:, 0:3] = d / dist[:, None]
            self.geo[:, 3:8] = v[:, [1, 3, 5, 7, 9]]
        h = np.maximum(dist.astype(np.float64) - 2.0, 1e-4)
        self.fe = np.zeros((self.E, FIN))
        if self.E:
            self.fe[:, 0] = 0.5 * np.log10(h)
            self.fe[:, 1:11] = 0.25 * np.sign(v) * np.log1p(np.abs(v))
        nlow = np.bincount(self.rows, minlength=N)
        nup = np.bincount(self.cols, minlength=N)
        trU = np.einsum('pii->p', Ad[:, 0:3, 0:3])
        trW = np.einsum('pii->p', Ad[:, 3:6, 3:6])
        self.fn = np.zeros((N, FIN))
        self.fn[:, 0] = 0.25 * nlow
        self.fn[:, 1] = 0.25 * nup
        self.fn[:, 2] = 0.25 * np.log1p(np.maximum(trU, 0) / 3.0)
        self.fn[:, 3] = 0.25 * np.log1p(np.maximum(trW, 0) / 3.0)
        # the pivot shift: lambda from 1, doubled while some A_PP + lambda S is not positive
        # definite (k_learned.hip lp_forward)
        if lam is None:
            lam = 1.0
            while not all(_is_pd(Ad[P] + lam * np.diag(S6)) for P in range(N)):
                lam *= 2.0
                assert lam <= 1e7
        self.lam = lam
        self.Dt = Ad + lam * np.einsum('c,cd->cd', S6, np.eye(6))[None]
        self.Dtinv = np.linalg.inv(self.Dt)
        # triangles of the lower pattern: (edge (P, Q), edge (P, R), edge (Q, R)) for every common
        # lower neighbour R of an edge's particles (the pair blocks' fill correction)
        low = {}
        for e in range(self.E):
            low.setdefault(int(self.rows[e]), {})[int(self.cols[e])] = e
        tri = []
        for e in range(self.E):
            lp_, lq = low.get(int(self.rows[e]), {}), low.get(int(self.cols[e]), {})
            for R, eq in lq.items():
                if R in lp_:
                    tri.append((e, lp_[R], eq, R))
        self.tri = np.array(tri, dtype=np.int64).reshape(-1, 4)
        self.levels_p = [np.zeros(N, dtype=np.int64) for _ in range(NPASS)]
        self.levels = self.levels_p[-1]
        self.Bk = basis(self.geo)

    # ---- network
    def _q(self, x, emulate):
        return bf16(x) if emulate else np.asarray(x, dtype=np.float64)

    def _agg(self, m):
        agg = np.zeros((self.N, m.shape[1]))
        np.add.at(agg, self.rows, m)
        np.add.at(agg, self.cols, m)
        return agg

    def forward(self, w, emulate=True):
        q = lambda x: self._q(x, emulate)
        Ls = [(q(Wl), b) for Wl, b in unpack(w)]
        c = {"xn0": q(self.fn), "xe0": q(self.fe)}
        c["zn0"] = c["xn0"] @ Ls[L_ENC_N][0] + Ls[L_ENC_N][1]
        c["ze0"] = c["xe0"] @ Ls[L_ENC_E][0] + Ls[L_ENC_E][1]
        h, e = [silu(c["zn0"])], [silu(c["ze0"])]
        for l in range(3):
            xm = q(np.concatenate([h[l][self.rows], h[l][self.cols], e[l]], axis=1))
            zm = xm @ Ls[L_MSG + 2 * l][0] + Ls[L_MSG + 2 * l][1]
            m = silu(zm)
            e.append(e[l] + m)
            agg = self._agg(m)
            xu = q(np.concatenate([h[l], agg], axis=1))
            zu = xu @ Ls[L_UPD + 2 * l][0] + Ls[L_UPD + 2 * l][1]
            h.append(h[l] + silu(zu))
            c["xm%d" % l], c["zm%d" % l], c["xu%d" % l], c["zu%d" % l] = xm, zm, xu, zu
        c["xm3"] = q(np.concatenate([h[3][self.rows], h[3][self.cols], e[3]], axis=1))
        c["oute"] = c["xm3"] @ Ls[L_HEAD_E][0] + Ls[L_HEAD_E][1]
        c["xh"] = q(h[3])
        c["outn"] = c["xh"] @ Ls[L_HEAD_N][0] + Ls[L_HEAD_N][1]
        c["h"], c["e"], c["Ls"] = h, e, Ls
        return c

    # ---- factor
    def assemble(self, oute, outn):
        """W [E,6,6], D, Dinv [N,6,6] and whether a network output is not finite.  The raw
        outputs enter bounded: multipliers 1 + tanh(c), fill weights 2 sigmoid(c + c0), pivot
        shift lambda e^a S (k_learned.hip lp_edge_W_kernel, lp_node_D_kernel).  Pivot
        safeguard: D_P = A_PP + lambda e^a S - s F_P with s the first of 1, 1/2, 1/4, 1/8
        for which D_P is positive definite (level 0 .. 3), else A_PP + lambda S (level 4);
        the levels are kept in self.levels for the backward pass."""
        W = np.einsum('ek,ekac->eac', 1.0 + np.tanh(oute[:, :6]), self.Bk) if self.E else np.zeros((0, 6, 6))
        B = self.Adiag.copy()
        B[:, [0, 1, 2], [0, 1, 2]] += self.lam * np.exp(outn[:, 0:1])
        B[:, [3, 4, 5], [3, 4, 5]] += self.lam * (4.0 / 3.0) * np.exp(outn[:, 1:2])
        g = gamma(oute[:, 6]) if self.E else np.zeros(0)
        failed = not (np.all(np.isfinite(outn[:, :2])) and np.all(np.isfinite(oute[:, :7])))

        def one_pass(Finv, levels):
            # the fill term uses the pair blocks A (the pivots do not depend on W)
            F = np.zeros_like(B)
            if self.E:
                T = np.einsum('eaq,eqr,ecr->eac', self.Ablk, Finv[self.cols], self.Ablk)
                np.add.at(F, self.rows, g[:, None, None] * T)
            D = np.empty_like(B)
            for P in range(self.N):
                levels[P] = LEVELS
                for j in range(LEVELS):
                    Dj = B[P] - 0.5 ** j * F[P]
                    Dj = np.tril(Dj) + np.tril(Dj, -1).T
                    if _is_pd(Dj):
                        levels[P], D[P] = j, Dj
                        break
                if levels[P] == LEVELS:
                    D[P] = self.Dt[P]
            return D

        # NPASS passes: the fill with the uncorrected pivots, then each with the previous pass's
        self.Dp, self.Dpinv = [], []
        Finv = self.Dtinv
        for p in range(NPASS):
            Dn = one_pass(Finv, self.levels_p[p])
            self.Dp.append(Dn)
            Finv = np.linalg.inv(Dn)
            self.Dpinv.append(Finv)
        # the pair blocks' one-step fill correction with the final pivots (k_learned.hip
        # lp_wfix_kernel): W_PQ -= sum_R A_PR D_R^-1 A_QR^T
        if len(self.tri):
            t = self.tri
            np.add.at(W, t[:, 0], -np.einsum('taq,tqr,tcr->tac', self.Ablk[t[:, 1]], Finv[t[:, 3]], self.Ablk[t[:, 2]]))
        return W, self.Dp[-1], self.Dpinv[-1], failed

    def lower(self, W, v):
        out = np.zeros((self.N, 6))
        if self.E:
            np.add.at(out, self.rows, np.einsum('eac,ec->ea', W, v[self.cols]))
        return out

    def upper(self, W, x):
        out = np.zeros((self.N, 6))
        if self.E:
            np.add.at(out, self.cols, np.einsum('eca,ec->ea', W, x[self.rows]))
        return out

    def apply_G(self, W, Dinv, d6, k=None):
        """G d with k block Jacobi iterations each way (d unpermuted, 6N)"""
        k = self.k if k is None else k
        t = np.asarray(d6, dtype=np.float64).reshape(self.N, 6)[self.perm]
        dm = lambda v: np.einsum('pac,pc->pa', Dinv, v)
        v = dm(t)
        for _ in range(k):
            v = dm(t - self.lower(W, v))
        vk, x = v, v
        for _ in range(k):
            x = vk - dm(self.upper(W, x))
        out = np.empty((self.N, 6))
        out[self.perm] = x
        return out.reshape(-1)

    def G_exact(self, W, D, d6):
        """((D + W) D^-1 (D + W)^T)^-1 d by dense solves (small N)"""
        N = self.N
        Lm = np.zeros((6 * N, 6 * N))
        for P in range(N):
            Lm[6 * P:6 * P + 6, 6 * P:6 * P + 6] = D[P]
        for e in range(self.E):
            P, Q = self.rows[e], self.cols[e]
            Lm[6 * P:6 * P + 6, 6 * Q:6 * Q + 6] = W[e]
        Dm = np.zeros_like(Lm)
        for P in range(N):
            Dm[6 * P:6 * P + 6, 6 * P:6 * P + 6] = D[P]
        Mm = Lm @ np.linalg.solve(Dm, Lm.T)
        t = np.asarray(d6, dtype=np.float64).reshape(N, 6)[self.perm].reshape(-1)
        y = np.linalg.solve(Mm, t).reshape(N, 6)
        out = np.empty((N, 6))
        out[self.perm] = y
        return out.reshape(-1)

    # ---- loss and gradient
    def loss(self, w, X, B, emulate=True):
        c = self.forward(w, emulate)
        W, D, Dinv, failed = self.assemble(c["oute"], c["outn"])
        return self._loss_from(W, Dinv, X, B)[0]

    def _loss_from(self, W, Dinv, X, B):
        N = self.N
        X, B = np.atleast_2d(X), np.atleast_2d(B)
        wsum = float((X.astype(np.float64) ** 2).sum())
        tot, gs = 0.0, []
        for xs, bs in zip(X, B):
            xs, bs = xs.astype(np.float64), bs.astype(np.float64)
            d = bs[:6 * N] - bs[11 * N:]
            g = self.apply_G(W, Dinv, d)
            y1 = bs[:6 * N] - xs[:6 * N]
            y2 = xs[11 * N:]
            tot += ((g - y1) ** 2).sum() + ((g - y2) ** 2).sum()
            gs.append(2.0 * ((g - y1) + (g - y2)) / wsum)
        return tot / wsum, gs

    def grad(self, w, X, B, emulate=True):
        """(loss, gradient over all parameters) by the same reverse sweep as the GPU"""
        N, k = self.N, self.k
        c = self.forward(w, emulate)
        W, D, Dinv, failed = self.assemble(c["oute"], c["outn"])
        loss, gbars = self._loss_from(W, Dinv, X, B)
        dm = lambda v: np.einsum('pac,pc->pa', Dinv, v)
        gW = np.zeros_like(W)
        gD = np.zeros((N, 6, 6))
        for xs, bs, gbar in zip(np.atleast_2d(X), np.atleast_2d(B), gbars):
            bs = bs.astype(np.float64)
            t = (bs[:6 * N] - bs[11 * N:]).reshape(N, 6)[self.perm]
            V, U, Xs, S = [dm(t)], [], [], []
            for j in range(k):
                U.append(t - self.lower(W, V[j]))
                V.append(dm(U[j]))
            Xs.append(V[k])
            for j in range(k):
                S.append(self.upper(W, Xs[j]))
                Xs.append(V[k] - dm(S[j]))
            Xb = [None] * (k + 1)
            Xb[k] = gbar.reshape(N, 6)[self.perm]
            Sb = [None] * k
            for j in range(k - 1, -1, -1):
                Sb[j] = -dm(Xb[j + 1])
                Xb[j] = self.lower(W, Sb[j])
            Vb = [None] * (k + 1)
            Vb[k] = sum(Xb)
            Ub = [None] * k
            for j in range(k - 1, -1, -1):
                Ub[j] = dm(Vb[j + 1])
                Vb[j] = -self.upper(W, Ub[j])
            for j in range(k):
                gW += -np.einsum('ea,ec->eac', Ub[j][self.rows], V[j][self.cols])
                gW += np.einsum('ea,ec->eac', Xs[j][self.rows], Sb[j][self.cols])
            gD += np.einsum('pa,pc->pac', Vb[0], t)
            for j in range(k):
                gD += np.einsum('pa,pc->pac', Vb[j + 1], U[j]) - np.einsum('pa,pc->pac', Xb[j + 1], S[j])
        # factor assembly backward
        Ssym = 0.5 * (gD + np.transpose(gD, (0, 2, 1)))
        Db = -np.einsum('paq,pqr,prc->pac', Dinv, Ssym, Dinv)
        return loss, self._assembly_backward(c, gW, Db, emulate)

    def factor_targets(self, orc):
        """IC(0)'s block LDL^T on the pattern: pivots L_PP L_PP^T and pair blocks L_PQ L_QQ^T"""
        Ld, rows, cols, Lo = orc.factor_blocks()
        assert np.array_equal(rows, self.rows) and np.array_equal(cols, self.cols)
        Dic = np.einsum('pac,pbc->pab', Ld, Ld)
        Wic = np.einsum('eac,ebc->eab', Lo, Ld[cols]) if self.E else np.zeros((0, 6, 6))
        return Dic, Wic

    def factor_loss(self, w, Dic, Wic, emulate=True, grad=False):
        """mean over particles of |D_P - D_P^IC|^2 / |D_P^IC|^2 plus mean over pairs of
        |W_e - W_e^IC|^2 / |A_e|^2 (k_learned.hip lp_pivot_target_kernel, lp_pair_target_kernel)"""
        c = self.forward(w, emulate)
        W, D, Dinv, failed = self.assemble(c["oute"], c["outn"])
        nD = (Dic ** 2).sum(axis=(1, 2))
        loss = float((((D - Dic) ** 2).sum(axis=(1, 2)) / nD).mean())
        gD = 2.0 * (D - Dic) / (self.N * nD[:, None, None])
        gW = np.zeros_like(W)
        if self.E:
            nA = (self.Ablk ** 2).sum(axis=(1, 2))
            loss += float((((W - Wic) ** 2).sum(axis=(1, 2)) / nA).mean())
            gW = 2.0 * (W - Wic) / (self.E * nA[:, None, None])
        if not grad:
            return loss
        Db = 0.5 * (gD + np.transpose(gD, (0, 2, 1)))
        return loss, self._assembly_backward(c, gW, Db, emulate)

    def _assembly_backward(self, c, gW, Db, emulate):
        N = self.N
        goutn = np.zeros((N, ON))
        outn = c["outn"]
        goute = np.zeros((self.E, OE))
        oute = c["oute"]

        def pass_backward(Db, levels, Finv):
            # one pivot pass: the shift and fill-weight gradients (none at level 4)
            live = levels < LEVELS
            sc = np.where(live, 0.5 ** np.minimum(levels, LEVELS - 1), 0.0)
            goutn[:, 0] += live * self.lam * np.exp(outn[:, 0]) * np.einsum('pii->p', Db[:, 0:3, 0:3])
            goutn[:, 1] += live * self.lam * (4.0 / 3.0) * np.exp(outn[:, 1]) * np.einsum('pii->p', Db[:, 3:6, 3:6])
            if self.E:
                T = np.einsum('eaq,eqr,ecr->eac', self.Ablk, Finv[self.cols], self.Ablk)
                goute[:, 6] += -sc[self.rows] * np.einsum('eac,eac->e', Db[self.rows], T) * gamma_d(oute[:, 6])
            return live, sc

        # the pair blocks' fill correction depends on the final pivots: -A_PR^T gW A_QR with
        # respect to D_R^-1, i.e. -D^-1 sym(.) D^-1 with respect to D_R
        if len(self.tri):
            t = self.tri
            G = np.zeros((N, 6, 6))
            np.add.at(G, t[:, 3], -np.einsum('tqa,tqr,trc->tac', self.Ablk[t[:, 1]], gW[t[:, 0]], self.Ablk[t[:, 2]]))
            Gs = 0.5 * (G + np.transpose(G, (0, 2, 1)))
            Dfi = self.Dpinv[-1]
            Db = Db - np.einsum('paq,pqr,prc->pac', Dfi, Gs, Dfi)
        # from the last pass to the first; each later pass depends on the earlier one's pivots
        # through its fill terms A D_Q^-1 A^T
        for p in range(NPASS - 1, -1, -1):
            Finv = self.Dpinv[p - 1] if p > 0 else self.Dtinv
            live, sc = pass_backward(Db, self.levels_p[p], Finv)
            if p == 0:
                break
            X = np.zeros((N, 6, 6))
            if self.E:
                Dbl = Db * live[:, None, None]
                w = -sc[self.rows] * gamma(oute[:, 6])
                np.add.at(X, self.cols, w[:, None, None] * np.einsum('eqa,eqr,erc->eac', self.Ablk, Dbl[self.rows],
                                                                       self.Ablk))
            Xs = 0.5 * (X + np.transpose(X, (0, 2, 1)))
            Db = -np.einsum('paq,pqr,prc->pac', Finv, Xs, Finv)
        if self.E:
            goute[:, :6] = np.einsum('eac,ekac->ek', gW, self.Bk) * (1.0 - np.tanh(oute[:, :6]) ** 2)
        return self._mlp_backward(c, goute, goutn, emulate)

    def _mlp_backward(self, c, goute, goutn, emulate):
        q = lambda x: self._q(x, emulate)
        Ls = c["Ls"]
        grads = [None] * len(LAYERS)

        def dense(l, X, gz):
            grads[l] = (X.T @ gz, gz.sum(axis=0))
            return q(gz) @ Ls[l][0].T

        gh = dense(L_HEAD_N, c["xh"], goutn)
        gx = dense(L_HEAD_E, c["xm3"], goute)
        gh = gh + self._scatter(gx)
        ge = gx[:, 2 * H:]
        for l in range(2, -1, -1):
            gzu = gh * silu_d(c["zu%d" % l])
            gxn = dense(L_UPD + 2 * l, c["xu%d" % l], gzu)
            gh = gh + gxn[:, :H]
            gagg = gxn[:, H:]
            gm = ge + gagg[self.rows] + gagg[self.cols]
            gzm = gm * silu_d(c["zm%d" % l])
            gx = dense(L_MSG + 2 * l, c["xm%d" % l], gzm)
            gh = gh + self._scatter(gx)
            ge = ge + gx[:, 2 * H:]
        grads[L_ENC_N] = (c["xn0"].T @ (gh * silu_d(c["zn0"])), (gh * silu_d(c["zn0"])).sum(axis=0))
        gze = ge * silu_d(c["ze0"])
        grads[L_ENC_E] = (c["xe0"].T @ gze, gze.sum(axis=0))
        offs, n = layer_offsets()
        g = np.zeros(n)
        for (wo, bo), (gw, gb), (ki, no, _) in zip(offs, grads, LAYERS):
            g[wo:wo + ki * no] = gw.reshape(-1)
            g[bo:bo + no] = gb
        return g

    def _scatter(self, gx):
        gh = np.zeros((self.N, H))
        np.add.at(gh, self.rows, gx[:, :H])
        np.add.at(gh, self.cols, gx[:, H:2 * H])
        return gh


def init_weights(seed=0, scale_out=0.0):
    """He-style initial weights (zero biases); the heads start at scale_out x that
    (0: the untrained factor is block symmetric Gauss-Seidel)"""
    rng = np.random.default_rng(seed)
    offs, n = layer_offsets()
    w = np.zeros(n, dtype=np.float32)
    for l, ((wo, bo), (ki, no, act)) in enumerate(zip(offs, LAYERS)):
        s = np.sqrt(2.0 / ki) * (scale_out if l in (L_HEAD_E, L_HEAD_N) else 1.0)
        w[wo:wo + ki * no] = (rng.standard_normal(ki * no) * s).astype(np.float32)
    return w

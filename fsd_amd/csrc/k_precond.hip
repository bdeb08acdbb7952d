// Near-field preconditioner: pruned neighbour list, R_FU^nf on the
// RCM-permuted particle graph in 6x6 blocks, block IC(0) and the two
// triangular solves of G = P^T L^-T L^-1 P.
//
// IC(0) on the reference's scalar pattern equals block IC(0) on the particle
// graph with dense 6x6 blocks: every stored pattern block is full (or, for an
// isolated particle, a diagonal block whose off-diagonal values are zero), so
// the scalar recurrences of csric02 reduce to block Cholesky without fill.
//
// Factorisation and solves are level-scheduled on the permuted particle DAG
// (one launch per level; the levels are built on the host with the RCM).
#include "fsd_internal.h"

// ----------------------------------------------------------- pruning
// Precondition_GetPrunedNneigh/GetPrunedNlist (Precondition.cu:70-242):
// neighbours with r^2 < rp^2 (r = x_i - x_j), self included, ascending.
__global__ void prune_kernel(const float4 *pos, int N, BoxF b, const int *nneigh, const int *nlist, int maxn,
                             float rp2, int *pn, int *plist, int *pless, int *overflow)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float4 pi = pos[i];
    int A[FSD_PSTRIDE];
    int cnt = 1;
    A[0] = i;
    for (int k = 0; k < nneigh[i]; ++k) {
        int j = nlist[(int64_t)i * maxn + k];
        float4 pj = pos[j];
        float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        box_min_image(b, dx, dy, dz);
        if (dist2_nc(dx, dy, dz) < rp2) {
            if (cnt < FSD_PSTRIDE) A[cnt] = j;
            ++cnt;
        }
    }
    if (cnt > FSD_PSTRIDE) { atomicMax(overflow, cnt); cnt = FSD_PSTRIDE; }
    for (int a = 1; a < cnt; ++a) {
        int key = A[a], q = a - 1;
        while (q >= 0 && A[q] > key) { A[q + 1] = A[q]; --q; }
        A[q + 1] = key;
    }
    int nl = 0;
    for (int a = 0; a < cnt; ++a) {
        plist[(int64_t)i * FSD_PSTRIDE + a] = A[a];
        nl += (A[a] < i);
    }
    pn[i] = cnt;
    pless[i] = nl;
}

int k_prune(fsd_ctx *ctx, const float4 *pos)
{
    const int N = ctx->N;
    FSD_HIP(hipMemsetAsync(ctx->d_flags + 2, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(prune_kernel, dim3(nblk(N, 128)), dim3(128), 0, ctx->stream, pos, N, ctx->box, ctx->d_nneigh,
                       ctx->d_nlist, ctx->maxn, ctx->p.rp * ctx->p.rp, ctx->d_pn, ctx->d_plist, ctx->d_pless,
                       ctx->d_flags + 2);
    FSD_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------------- block values
// Precondition_RFU_kernel (Precondition.cu:530-836) in block form: for the
// permuted particle P (original i = perm[P]) the diagonal block accumulates
// the self terms over its pruned neighbours and each lower neighbour Q < P
// gets the pair block [[A, B], [B, C]] (row particle i, column particle j).
__global__ void blocks_kernel(const float4 *pos, int N, BoxF b, const int *pn, const int *plist, const int *perm,
                              const int *iperm, const int *lrow, const int *lcol, LubTable t, float rp2,
                              float *Adiag, float *Ablk)
{
    int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= N) return;
    const int i = perm[P];
    float4 pi = pos[i];
    float S[36];
    for (int q = 0; q < 36; ++q) S[q] = 0.f;
    const int n = pn[i];
    for (int s = 0; s < n; ++s) {
        const int j = plist[(int64_t)i * FSD_PSTRIDE + s];
        if (j == i) continue;
        float4 pj = pos[j];
        float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        box_min_image(b, dx, dy, dz);
        const float d2 = dist2_nc(dx, dy, dz);
        if (!(d2 < rp2)) continue;
        const float dist = sqrtf(d2);
        float v[10];
        if ((double)dist <= 2.0 + (double)t.table_min) {
            for (int c = 0; c < 10; ++c) v[c] = t.vals[c];     // contact clamp: table row 0
        } else {
            lub_lookup(t, dist, 0, 10, -1, v);
        }
        const float XA11 = v[0], XA12 = v[1], YA11 = v[2], YA12 = v[3], YB11 = v[4], YB12 = v[5];
        const float XC11 = v[6], XC12 = v[7], YC11 = v[8], YC12 = v[9];
        const float r[3] = {dx / dist, dy / dist, dz / dist};
        const float E[3][3] = {{0.f, r[2], -r[1]}, {-r[2], 0.f, r[0]}, {r[1], -r[0], 0.f}};
        const int Q = iperm[j];
        float *blk = nullptr;
        if (Q < P) {
            int lo = lrow[P], hi = lrow[P + 1] - 1;
            while (lo < hi) { int mid = (lo + hi) >> 1; if (lcol[mid] < Q) lo = mid + 1; else hi = mid; }
            if (lo < lrow[P + 1] && lcol[lo] == Q) blk = Ablk + (int64_t)lo * 36;
        }
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) {
                const float rr = r[a] * r[c];
                const float imrr = ((a == c) ? 1.0f : 0.0f) - rr;
                S[a * 6 + c] += XA11 * rr + YA11 * imrr;
                S[a * 6 + 3 + c] += -(YB11 * E[a][c]);
                S[(3 + a) * 6 + c] += YB11 * E[a][c];
                S[(3 + a) * 6 + 3 + c] += XC11 * rr + YC11 * imrr;
                if (blk) {
                    const float Bn = YB12 * E[a][c];
                    blk[a * 6 + c] = XA12 * rr + YA12 * imrr;
                    blk[a * 6 + 3 + c] = Bn;
                    blk[(3 + a) * 6 + c] = Bn;
                    blk[(3 + a) * 6 + 3 + c] = XC12 * rr + YC12 * imrr;
                }
            }
    }
    for (int q = 0; q < 36; ++q) Adiag[(int64_t)P * 36 + q] = S[q];
}

int k_build_blocks(fsd_ctx *ctx, const float4 *pos)
{
    LubTable t{ctx->d_res_dist, ctx->d_res_vals, ctx->table_min, ctx->table_dr};
    hipLaunchKernelGGL(blocks_kernel, dim3(nblk(ctx->N, 128)), dim3(128), 0, ctx->stream, pos, ctx->N, ctx->box,
                       ctx->d_pn, ctx->d_plist, ctx->d_perm, ctx->d_iperm, ctx->d_lrow, ctx->d_lcol, t,
                       ctx->p.rp * ctx->p.rp, ctx->d_Adiag, ctx->d_Ablk);
    FSD_HIP(hipGetLastError());
    return 0;
}

// -------------------------------------------------------------- IC(0)
// Level-scheduled: the host groups the permuted particles into levels of the
// lower-triangular particle DAG (a particle's level is one more than the
// highest level among its lower neighbours); one launch per level, one
// wavefront per particle, lane (a*6+b) owning entry (a,b).  Kernel
// boundaries order the levels, so no particle ever waits on another.
__global__ void __launch_bounds__(64) ichol_level_kernel(const int *items, const int *lrow, const int *lcol,
                                                         const float *Ablk, const float *Adiag, float lambda,
                                                         float *Lblk, float *Ldiag, int *fail)
{
    __shared__ float T[36];
    __shared__ float Lq[36];
    __shared__ float Dm[36];
    const int lane = threadIdx.x;
    const int a = lane / 6, bcol = lane % 6;
    const bool ent = lane < 36;
    const int P = items[blockIdx.x];
    const int beg = lrow[P], end = lrow[P + 1];
    float dacc = 0.f;  // sum_Q L_PQ L_PQ^T for entry (a, bcol)
    for (int e = beg; e < end; ++e) {
        const int Q = lcol[e];
        // T = A_PQ - sum_{R < Q, R in lower(P) and lower(Q)} L_PR L_QR^T
        if (ent) {
            float tv = Ablk[(int64_t)e * 36 + lane];
            int p1 = beg, p2 = lrow[Q];
            const int e1 = e, e2 = lrow[Q + 1];
            while (p1 < e1 && p2 < e2) {
                const int c1 = lcol[p1], c2 = lcol[p2];
                if (c1 == c2) {
                    const float *x = Lblk + (int64_t)p1 * 36 + a * 6;
                    const float *y = Lblk + (int64_t)p2 * 36 + bcol * 6;
                    float sacc = 0.f;
                    for (int c = 0; c < 6; ++c) sacc += x[c] * y[c];
                    tv -= sacc;
                    ++p1; ++p2;
                } else if (c1 < c2) ++p1;
                else ++p2;
            }
            T[lane] = tv;
            Lq[lane] = Ldiag[(int64_t)Q * 36 + lane];
        }
        __syncthreads();
        // X L_QQ^T = T, column by column
        for (int c = 0; c < 6; ++c) {
            if (ent && bcol == c) {
                float v = T[a * 6 + c];
                for (int q = 0; q < c; ++q) v -= T[a * 6 + q] * Lq[c * 6 + q];
                T[a * 6 + c] = v / Lq[c * 6 + c];
            }
            __syncthreads();
        }
        if (ent) {
            Lblk[(int64_t)e * 36 + lane] = T[lane];
            float sacc = 0.f;
            for (int c = 0; c < 6; ++c) sacc += T[a * 6 + c] * T[bcol * 6 + c];
            dacc += sacc;
        }
        __syncthreads();
    }
    if (ent) {
        float v = Adiag[(int64_t)P * 36 + lane] - dacc;
        if (a == bcol) v += lambda * ((a < 3) ? 1.0f : 1.33333333f);
        Dm[lane] = v;
    }
    __syncthreads();
    if (lane == 0) {
        bool bad = false;
        for (int c = 0; c < 6; ++c) {
            float sd = Dm[c * 6 + c];
            for (int q = 0; q < c; ++q) sd -= Dm[c * 6 + q] * Dm[c * 6 + q];
            if (!(sd > 0.f)) { bad = true; sd = 1.f; }
            const float d = sqrtf(sd);
            Dm[c * 6 + c] = d;
            for (int r = c + 1; r < 6; ++r) {
                float v = Dm[r * 6 + c];
                for (int q = 0; q < c; ++q) v -= Dm[r * 6 + q] * Dm[c * 6 + q];
                Dm[r * 6 + c] = v / d;
            }
        }
        for (int r = 0; r < 6; ++r)
            for (int c = r + 1; c < 6; ++c) Dm[r * 6 + c] = 0.f;
        if (bad) atomicMax(fail, 1);
    }
    __syncthreads();
    if (ent) Ldiag[(int64_t)P * 36 + lane] = Dm[lane];
}

int k_ichol(fsd_ctx *ctx, float lambda, int *failed)
{
    hipEvent_t ev;
    prof_begin(ctx, PROF_ICHOL, &ev);
    FSD_HIP(hipMemsetAsync(ctx->d_flags + 2, 0, sizeof(int), ctx->stream));
    for (int l = 0; l < ctx->nlev_f; ++l) {
        const int b = ctx->h_lev_f_ptr[l], n = ctx->h_lev_f_ptr[l + 1] - b;
        hipLaunchKernelGGL(ichol_level_kernel, dim3(n), dim3(64), 0, ctx->stream, ctx->d_lev_f + b, ctx->d_lrow,
                           ctx->d_lcol, ctx->d_Ablk, ctx->d_Adiag, lambda, ctx->d_Lblk, ctx->d_Ldiag,
                           ctx->d_flags + 2);
    }
    FSD_HIP(hipGetLastError());
    FSD_HIP(hipMemcpyAsync(ctx->h_flags + 2, ctx->d_flags + 2, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FSD_HIP(hipStreamSynchronize(ctx->stream));
    prof_end(ctx, PROF_ICHOL, ev);
    *failed = ctx->h_flags[2];
    return 0;
}

// --------------------------------------------------- triangular solves
// Forward level: y_P = L_PP^-1 (b_perm[P] - sum_{Q<P} L_PQ y_Q) for nrhs
// vectors; lanes 6*r + a (r < nrhs, a < 6) own row a of right-hand side r.
__global__ void __launch_bounds__(64) trsv_fwd_level_kernel(const int *items, int N, int nrhs, const int *lrow,
                                                            const int *lcol, const float *Lblk, const float *Ldiag,
                                                            const int *perm, const float *b0, const float *b1,
                                                            float *y)
{
    __shared__ float s_v[12];
    const int lane = threadIdx.x;
    const int r = lane / 6, a = lane % 6;
    const bool act = lane < 6 * nrhs;
    const int P = items[blockIdx.x];
    const int beg = lrow[P], end = lrow[P + 1];
    if (act) {
        const float *bsrc = (r == 0) ? b0 : b1;
        float v = bsrc[6 * (int64_t)perm[P] + a];
        for (int e = beg; e < end; ++e) {
            const int Q = lcol[e];
            const float *L = Lblk + (int64_t)e * 36 + a * 6;
            const float *yq = y + (int64_t)r * 6 * N + 6 * (int64_t)Q;
            for (int c = 0; c < 6; ++c) v -= L[c] * yq[c];
        }
        s_v[lane] = v;
    }
    __syncthreads();
    if (lane < nrhs) {
        const float *D = Ldiag + (int64_t)P * 36;
        float *v = s_v + 6 * lane;
        for (int c = 0; c < 6; ++c) {
            float sv = v[c];
            for (int q = 0; q < c; ++q) sv -= D[c * 6 + q] * v[q];
            v[c] = sv / D[c * 6 + c];
        }
    }
    __syncthreads();
    if (act) y[(int64_t)r * 6 * N + 6 * (int64_t)P + a] = s_v[lane];
}

// Backward level: x_P = L_PP^-T (y_P - sum_{Q>P} L_QP^T x_Q); results are also
// written un-permuted: out[6 perm[P] + a].
__global__ void __launch_bounds__(64) trsv_bwd_level_kernel(const int *items, int N, int nrhs, const int *urow,
                                                            const int *ucol, const int *upos, const float *Lblk,
                                                            const float *Ldiag, const int *perm, const float *y,
                                                            float *x, float *out0, float *out1)
{
    __shared__ float s_v[12];
    const int lane = threadIdx.x;
    const int r = lane / 6, a = lane % 6;
    const bool act = lane < 6 * nrhs;
    const int P = items[blockIdx.x];
    const int beg = urow[P], end = urow[P + 1];
    if (act) {
        float v = y[(int64_t)r * 6 * N + 6 * (int64_t)P + a];
        for (int e = beg; e < end; ++e) {
            const int Q = ucol[e];
            const float *L = Lblk + (int64_t)upos[e] * 36;   // block (Q, P)
            const float *xq = x + (int64_t)r * 6 * N + 6 * (int64_t)Q;
            for (int c = 0; c < 6; ++c) v -= L[c * 6 + a] * xq[c];
        }
        s_v[lane] = v;
    }
    __syncthreads();
    if (lane < nrhs) {
        const float *D = Ldiag + (int64_t)P * 36;
        float *v = s_v + 6 * lane;
        for (int c = 5; c >= 0; --c) {
            float sv = v[c];
            for (int q = c + 1; q < 6; ++q) sv -= D[q * 6 + c] * v[q];
            v[c] = sv / D[c * 6 + c];
        }
    }
    __syncthreads();
    if (act) {
        const float val = s_v[lane];
        x[(int64_t)r * 6 * N + 6 * (int64_t)P + a] = val;
        float *o = (r == 0) ? out0 : out1;
        o[6 * (int64_t)perm[P] + a] = val;
    }
}

// y_r = G b_r for r < n_rhs (b1/y1 used when n_rhs == 2)
int k_precond_apply2(fsd_ctx *ctx, const float *b0, const float *b1, float *y0, float *y1, int n_rhs)
{
    const int N = ctx->N;
    hipEvent_t ev;
    prof_begin(ctx, PROF_TRSV, &ev);
    for (int l = 0; l < ctx->nlev_f; ++l) {
        const int b = ctx->h_lev_f_ptr[l], n = ctx->h_lev_f_ptr[l + 1] - b;
        hipLaunchKernelGGL(trsv_fwd_level_kernel, dim3(n), dim3(64), 0, ctx->stream, ctx->d_lev_f + b, N, n_rhs,
                           ctx->d_lrow, ctx->d_lcol, ctx->d_Lblk, ctx->d_Ldiag, ctx->d_perm, b0, b1, ctx->d_tmp6a);
    }
    for (int l = 0; l < ctx->nlev_b; ++l) {
        const int b = ctx->h_lev_b_ptr[l], n = ctx->h_lev_b_ptr[l + 1] - b;
        hipLaunchKernelGGL(trsv_bwd_level_kernel, dim3(n), dim3(64), 0, ctx->stream, ctx->d_lev_b + b, N, n_rhs,
                           ctx->d_urow, ctx->d_ucol, ctx->d_upos, ctx->d_Lblk, ctx->d_Ldiag, ctx->d_perm,
                           ctx->d_tmp6a, ctx->d_tmp6b, y0, y1);
    }
    FSD_HIP(hipGetLastError());
    prof_end(ctx, PROF_TRSV, ev);
    return 0;
}

// Near-field preconditioner: pruned neighbour list, R_FU^nf on the
// RCM-permuted particle graph in 6x6 blocks, block IC(0) and the two
// triangular solves of G = P^T L^-T L^-1 P.
//
// IC(0) on the reference's scalar pattern equals block IC(0) on the particle
// graph with dense 6x6 blocks: every stored pattern block is full (or, for an
// isolated particle, a diagonal block whose off-diagonal values are zero), so
// the scalar recurrences of csric02 reduce to block Cholesky without fill.
//
// The factorisation is level-scheduled on the permuted particle DAG (one
// launch per level; the levels are built on the host with the RCM).  The
// solves run both sweeps in one single-workgroup launch on a static
// schedule (sweep_plan.h), or one launch per level where LDS is too small.
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
//
// The same kernel runs the parallel fixed-point form (Chow & Patel, "Fine-grained
// parallel incomplete LU factorization", SIAM J. Sci. Comput. 2015): items = null
// puts every particle in one launch, and the rows Q < P are read from the previous
// sweep's factor (Lblk_q, Ldiag_q) while row P's own new blocks are read back
// from the output.  Its fixed point is the IC(0) factor of the level recurrence.
__global__ void __launch_bounds__(64) ichol_level_kernel(const int *items, const int *lrow, const int *lcol,
                                                         const float *Ablk, const float *Adiag, float lambda,
                                                         float *Lblk, float *Ldiag, const float *Lblk_q,
                                                         const float *Ldiag_q, int *fail)
{
    __shared__ float T[36];
    __shared__ float Lq[36];
    __shared__ float Dm[36];
    const int lane = threadIdx.x;
    const int a = lane / 6, bcol = lane % 6;
    const bool ent = lane < 36;
    const int P = items ? items[blockIdx.x] : (int)blockIdx.x;
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
                    const float *y = Lblk_q + (int64_t)p2 * 36 + bcol * 6;
                    float sacc = 0.f;
                    for (int c = 0; c < 6; ++c) sacc += x[c] * y[c];
                    tv -= sacc;
                    ++p1; ++p2;
                } else if (c1 < c2) ++p1;
                else ++p2;
            }
            T[lane] = tv;
            Lq[lane] = Ldiag_q[(int64_t)Q * 36 + lane];
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
    for (int l = 0; l < ctx->pch.nlev_f; ++l) {
        const int b = ctx->pch.lev_f_ptr[l], n = ctx->pch.lev_f_ptr[l + 1] - b;
        hipLaunchKernelGGL(ichol_level_kernel, dim3(n), dim3(64), 0, ctx->stream, ctx->d_lev_f + b, ctx->d_lrow,
                           ctx->d_lcol, ctx->d_Ablk, ctx->d_Adiag, lambda, ctx->d_Lblk, ctx->d_Ldiag, ctx->d_Lblk,
                           ctx->d_Ldiag, ctx->d_flags + 2);
    }
    FSD_HIP(hipGetLastError());
    FSD_HIP(hipMemcpyAsync(ctx->h_flags + 2, ctx->d_flags + 2, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    FSD_HIP(hipStreamSynchronize(ctx->stream));
    prof_end(ctx, PROF_ICHOL, ev);
    *failed = ctx->h_flags[2];
    return 0;
}

// start of the parallel sweeps: L_PP = chol(A_PP + lambda shift), no blocks below the diagonal
__global__ void ichol_start_kernel(int N, const float *Adiag, float lambda, float *Ldiag, int *fail)
{
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= N) return;
    float D[36];
    for (int q = 0; q < 36; ++q) D[q] = Adiag[(int64_t)P * 36 + q];
    for (int c = 0; c < 6; ++c) D[7 * c] += lambda * ((c < 3) ? 1.0f : 1.33333333f);
    bool bad = false;
    for (int c = 0; c < 6; ++c) {
        float sd = D[c * 6 + c];
        for (int q = 0; q < c; ++q) sd -= D[c * 6 + q] * D[c * 6 + q];
        if (!(sd > 0.f)) { bad = true; sd = 1.f; }
        const float d = sqrtf(sd);
        D[c * 6 + c] = d;
        for (int r = c + 1; r < 6; ++r) {
            float v = D[r * 6 + c];
            for (int q = 0; q < c; ++q) v -= D[r * 6 + q] * D[c * 6 + q];
            D[r * 6 + c] = v / d;
        }
    }
    for (int r = 0; r < 6; ++r)
        for (int c = r + 1; c < 6; ++c) D[r * 6 + c] = 0.f;
    for (int q = 0; q < 36; ++q) Ldiag[(int64_t)P * 36 + q] = D[q];
    if (bad) atomicMax(fail, 1);
}

__global__ void zero_kernel(int64_t n, float *x)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.f;
}

// IC(0) by `sweeps` parallel fixed-point sweeps; the factor ends in (d_Lblk, d_Ldiag).
// A non-positive pivot of the last sweep fails the attempt (the caller doubles lambda).
int k_ichol_sweeps(fsd_ctx *ctx, float lambda, int sweeps, int *failed)
{
    const int N = ctx->N;
    const int64_t nl = ctx->pch.nlow;
    hipEvent_t ev;
    prof_begin(ctx, PROF_ICHOL, &ev);
    // buffers alternate so that the last sweep writes d_Lblk / d_Ldiag
    float *bl[2] = {ctx->d_Lblk, ctx->d_Lblk2}, *bd[2] = {ctx->d_Ldiag, ctx->d_Ldiag2};
    int cur = sweeps & 1;   // buffer holding the start
    FSD_HIP(hipMemsetAsync(ctx->d_flags + 2, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(ichol_start_kernel, dim3(nblk(N, 128)), dim3(128), 0, ctx->stream, N, ctx->d_Adiag, lambda,
                       bd[cur], ctx->d_flags + 2);
    if (nl) hipLaunchKernelGGL(zero_kernel, dim3(nblk(36 * nl, 256)), dim3(256), 0, ctx->stream, 36 * nl, bl[cur]);
    for (int s = 0; s < sweeps; ++s) {
        FSD_HIP(hipMemsetAsync(ctx->d_flags + 2, 0, sizeof(int), ctx->stream));   // pivots of the last sweep decide
        hipLaunchKernelGGL(ichol_level_kernel, dim3(N), dim3(64), 0, ctx->stream, (const int *)nullptr, ctx->d_lrow,
                           ctx->d_lcol, ctx->d_Ablk, ctx->d_Adiag, lambda, bl[cur ^ 1], bd[cur ^ 1], bl[cur],
                           bd[cur], ctx->d_flags + 2);
        cur ^= 1;
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

// ------------------------------------- single-workgroup sweeps
// Each sweep of one application runs in one workgroup of 1024 threads, so no
// workgroup ever waits on another.  The diagonal blocks are folded out of
// the recurrences (Z_P = L_PP^-1, lower triangular):
//   forward   y_P = Z_P b_P - sum_Q (Z_P L_PQ) y_Q
//   backward  u_P = y_P - sum_Q (Z_Q L_QP)^T u_Q,   x_P = Z_P^T u_P
// so a step is one 6x6 block product per lane, a sum over the row's lanes
// and a subtraction; Z b, the reordering of y between the sweeps and Z^T u
// are separate all-parallel kernels.
//
// The schedule (sweep_plan.h) gives each lane one dependency block per step.
// Results of the last steps' rows are read from an LDS ring, older ones from
// global memory; every operand of step s + 1 (blocks, inputs, the values of
// dependencies old enough to be read from global memory) is fetched while
// step s runs.  A workgroup barrier on gfx950 waits for LDS only (the memory
// model orders global memory inside a workgroup without waiting for stores),
// so stores stay in flight across steps.
struct TaskDev {
    int nsteps;
    int nrows;              // rows (= particles): stride of the per-row arrays
    const int2 *info;       // per lane: src, own
    const float4 *blk;      // [step][9][1024]: the lane's 6x6 block, row-major
    const float4 *inp;      // [3][nrows]: the rows' inputs, right-hand sides 0 and 1 (6 floats each)
    float4 *res;            // [3][nrows]: the rows' results
};

__device__ __forceinline__ uint32_t tp_idx(uint32_t g, uint32_t i, uint32_t n)
{
    return ((g >> 10) * n + i) * (uint32_t)TP_LANES + (g & (TP_LANES - 1));
}

// operands of one step for this lane (values of 2 right-hand sides as 3 float4:
// (a0 a1 a2 a3) (a4 a5 b0 b1) (b2 b3 b4 b5))
template <int NRHS>
__device__ __forceinline__ void tp_fetch(const TaskDev &S, uint32_t s, uint32_t t, int2 inf, float *m, float4 *gx,
                                         float4 *in)
{
    constexpr int NW = NRHS == 2 ? 3 : 2;
    const uint32_t nr = (uint32_t)S.nrows;
    if (inf.x >= 0) {
        const float4 *b = S.blk + (size_t)s * 9 * TP_LANES + t;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const float4 v = b[i * TP_LANES];
            m[4 * i] = v.x; m[4 * i + 1] = v.y; m[4 * i + 2] = v.z; m[4 * i + 3] = v.w;
        }
        if (inf.x & TP_GLOBAL) {
            const uint32_t k = (uint32_t)(inf.x & ~TP_GLOBAL);
#pragma unroll
            for (int i = 0; i < NW; ++i) gx[i] = S.res[i * nr + k];
        }
    }
    if (inf.y >= 0) {
        const uint32_t k = (uint32_t)(inf.y & ((1 << TP_ND_SHIFT) - 1));
#pragma unroll
        for (int i = 0; i < NW; ++i) in[i] = S.inp[i * nr + k];
    }
}

template <int NRHS>
__device__ __forceinline__ void tp_unpack(const float4 *w, float *a, float *b)
{
    a[0] = w[0].x; a[1] = w[0].y; a[2] = w[0].z; a[3] = w[0].w; a[4] = w[1].x; a[5] = w[1].y;
    if (NRHS == 2) { b[0] = w[1].z; b[1] = w[1].w; b[2] = w[2].x; b[3] = w[2].y; b[4] = w[2].z; b[5] = w[2].w; }
}

template <int NRHS>
__device__ __forceinline__ void tp_pack(const float *a, const float *b, float4 *w)
{
    w[0] = make_float4(a[0], a[1], a[2], a[3]);
    w[1] = make_float4(a[4], a[5], NRHS == 2 ? b[0] : 0.f, NRHS == 2 ? b[1] : 0.f);
    if (NRHS == 2) w[2] = make_float4(b[2], b[3], b[4], b[5]);
}

// one sweep in one workgroup of 1024 threads; dynamic LDS: ring [3][TP_RING]
// and the owners' partial sums [3][1024] (float4)
template <int NRHS>
__global__ void __launch_bounds__(TP_LANES) tp_sweep_kernel(TaskDev S)
{
    constexpr int NW = NRHS == 2 ? 3 : 2;   // float4 per row value
    extern __shared__ float4 tp_lds[];
    float4 *ring = tp_lds;
    float4 *red = tp_lds + 3 * TP_RING;
    const uint32_t t = threadIdx.x;
    const uint32_t nr = (uint32_t)S.nrows;
    if (S.nsteps <= 0) return;
    // positions are fetched two steps ahead; past the last step the last one is re-read (and not used)
    const int last = S.nsteps - 1;
    int2 cur = S.info[t];
    int2 nxt = S.info[(size_t)min(1, last) * TP_LANES + t];
    float m[36];
    float4 gx[3], in[3];
    tp_fetch<NRHS>(S, 0, t, cur, m, gx, in);
    // this lane's result of the previous step (owner), stored during the next step
    float4 pw[3];
    int pk = -1;
    for (int s = 0; s < S.nsteps; ++s) {
        // everything in flight is needed now: this step's operands, the next step's
        // positions and (issued before them) the stores of two steps back
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
        const bool has = cur.x >= 0, own = cur.y >= 0;
        float pa[6], pb[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) { pa[a] = 0.f; pb[a] = 0.f; }
        if (has) {
            float4 xv[3];
            if (cur.x & TP_GLOBAL) {
#pragma unroll
                for (int i = 0; i < NW; ++i) xv[i] = gx[i];
            } else {
#pragma unroll
                for (int i = 0; i < NW; ++i) xv[i] = ring[i * TP_RING + cur.x];
            }
            float xa[6], xb[6];
            tp_unpack<NRHS>(xv, xa, xb);
#pragma unroll
            for (int a = 0; a < 6; ++a) {
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    pa[a] += m[6 * a + c] * xa[c];
                    if (NRHS == 2) pb[a] += m[6 * a + c] * xb[c];
                }
            }
        }
        float wa[6], wb[6];
        if (own) {
            tp_unpack<NRHS>(in, wa, wb);
#pragma unroll
            for (int a = 0; a < 6; ++a) { wa[a] -= pa[a]; if (NRHS == 2) wb[a] -= pb[a]; }
        }
        const int k = cur.y & ((1 << TP_ND_SHIFT) - 1), nd = cur.y >> TP_ND_SHIFT;
        // global stores of the previous step, then the operands of step s + 1 (positions
        // fetched one step earlier) and the positions of step s + 2: waiting for the
        // operands never waits for a store issued after them
        if (pk >= 0) {
#pragma unroll
            for (int i = 0; i < NW; ++i) S.res[i * nr + (uint32_t)pk] = pw[i];
        }
        if (s + 1 < S.nsteps) tp_fetch<NRHS>(S, (uint32_t)s + 1, t, nxt, m, gx, in);
        const int2 n2 = S.info[(size_t)min(s + 2, last) * TP_LANES + t];
        // the row's other lanes hand their products to its owner (same wave)
        if (has && !own) {
            float4 w[3];
            tp_pack<NRHS>(pa, pb, w);
#pragma unroll
            for (int i = 0; i < NW; ++i) red[i * TP_LANES + t] = w[i];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
        if (own) {
            // partial sums of the row's other lanes, four lanes per batch of loads
            for (int d0 = 1; d0 < nd; d0 += 4) {
                float4 w[4][3];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t lu = t + (uint32_t)min(d0 + u, nd - 1);   // in-row lane (repeats are masked below)
#pragma unroll
                    for (int i = 0; i < NW; ++i) w[u][i] = red[i * TP_LANES + lu];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (d0 + u < nd) {
                        float qa[6], qb[6];
                        tp_unpack<NRHS>(w[u], qa, qb);
#pragma unroll
                        for (int a = 0; a < 6; ++a) { wa[a] -= qa[a]; if (NRHS == 2) wb[a] -= qb[a]; }
                    }
                }
            }
            tp_pack<NRHS>(wa, wb, pw);
#pragma unroll
            for (int i = 0; i < NW; ++i) ring[i * TP_RING + (k & (TP_RING - 1))] = pw[i];
        }
        pk = own ? k : -1;
        __syncthreads();
        cur = nxt;
        nxt = n2;
    }
    if (pk >= 0) {
#pragma unroll
        for (int i = 0; i < NW; ++i) S.res[i * nr + (uint32_t)pk] = pw[i];
    }
}

// One right-hand side per workgroup: the launch has one workgroup per
// right-hand side (r = blockIdx.x), on different CUs, so each CU does half the
// products and reductions of the joint kernel above and reads the same
// blocks.  Row values of right-hand side r are the 6 floats (a0..a5) of the
// [3][nrows] float4 layout: r = 0 at (f4[0], f4[1].xy), r = 1 at
// (f4[1].zw, f4[2]).
template <int r>
__device__ __forceinline__ void row_ld(const float4 *A, uint32_t nr, uint32_t k, float *v)
{
    if constexpr (r == 0) {
        const float4 a = A[k];
        const float2 b = reinterpret_cast<const float2 *>(A + nr + k)[0];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y;
    } else {
        const float2 a = reinterpret_cast<const float2 *>(A + nr + k)[1];
        const float4 b = A[2 * nr + k];
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y; v[4] = b.z; v[5] = b.w;
    }
}

template <int r>
__device__ __forceinline__ void row_st(float4 *A, uint32_t nr, uint32_t k, const float *v)
{
    if constexpr (r == 0) {
        A[k] = make_float4(v[0], v[1], v[2], v[3]);
        reinterpret_cast<float2 *>(A + nr + k)[0] = make_float2(v[4], v[5]);
    } else {
        reinterpret_cast<float2 *>(A + nr + k)[1] = make_float2(v[0], v[1]);
        A[2 * nr + k] = make_float4(v[2], v[3], v[4], v[5]);
    }
}

// LDS: ring [TP_RING] float4 + [TP_RING] float2, partial sums [1024] float4 + [1024] float2
#define TP1_LDS (TP_RING * 24 + TP_LANES * 24)
// The sweep of right-hand side r (compile-time, so its row loads and stores are whole float4 /
// float2 accesses); the kernel below calls it for r = blockIdx.x.
template <int r>
__device__ __forceinline__ void tp_sweep1_body(const TaskDev &S, float4 *tp1_lds)
{
    float4 *ring4 = tp1_lds;
    float2 *ring2 = reinterpret_cast<float2 *>(tp1_lds + TP_RING);
    float4 *red4 = reinterpret_cast<float4 *>(ring2 + TP_RING);
    float2 *red2 = reinterpret_cast<float2 *>(red4 + TP_LANES);
    const uint32_t t = threadIdx.x;
    const uint32_t nr = (uint32_t)S.nrows;
    if (S.nsteps <= 0) return;
    const int last = S.nsteps - 1;
    int2 cur = S.info[t];
    int2 nxt = S.info[(size_t)min(1, last) * TP_LANES + t];
    float m[36], gx[6], in[6];
    auto fetch = [&](uint32_t s, int2 inf) {
        if (inf.x >= 0) {
            const float4 *b = S.blk + (size_t)s * 9 * TP_LANES + t;
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                const float4 v = b[i * TP_LANES];
                m[4 * i] = v.x; m[4 * i + 1] = v.y; m[4 * i + 2] = v.z; m[4 * i + 3] = v.w;
            }
            if (inf.x & TP_GLOBAL) row_ld<r>(S.res, nr, (uint32_t)(inf.x & ~TP_GLOBAL), gx);
        }
        if (inf.y >= 0) row_ld<r>(S.inp, nr, (uint32_t)(inf.y & ((1 << TP_ND_SHIFT) - 1)), in);
    };
    fetch(0, cur);
    float pw[6];
    int pk = -1;
    for (int s = 0; s < S.nsteps; ++s) {
        // the dependency's value from the LDS ring is read first (its slot was written before the
        // barrier that ended the previous step; lanes whose dependency is global or absent read a
        // valid slot whose value is not used), so its latency overlaps the wait below
        const uint32_t rs = (uint32_t)cur.x & (TP_RING - 1);
        const float4 ru = ring4[rs];
        const float2 rv = ring2[rs];
        // everything in flight is needed now: this step's operands, the next step's
        // positions and (issued before them) the stores of two steps back
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
        const bool has = cur.x >= 0, own = cur.y >= 0;
        float p[6];
#pragma unroll
        for (int a = 0; a < 6; ++a) p[a] = 0.f;
        if (has) {
            const bool glob = (cur.x & TP_GLOBAL) != 0;
            const float x[6] = {glob ? gx[0] : ru.x, glob ? gx[1] : ru.y, glob ? gx[2] : ru.z,
                                glob ? gx[3] : ru.w, glob ? gx[4] : rv.x, glob ? gx[5] : rv.y};
#pragma unroll
            for (int a = 0; a < 6; ++a)
#pragma unroll
                for (int c = 0; c < 6; ++c) p[a] += m[6 * a + c] * x[c];
        }
        float w[6];
        if (own) {
#pragma unroll
            for (int a = 0; a < 6; ++a) w[a] = in[a] - p[a];
        }
        const int k = cur.y & ((1 << TP_ND_SHIFT) - 1), nd = cur.y >> TP_ND_SHIFT;
        // global store of the previous step, then the operands of step s + 1 and positions of s + 2
        if (pk >= 0) row_st<r>(S.res, nr, (uint32_t)pk, pw);
        if (s + 1 < S.nsteps) fetch((uint32_t)s + 1, nxt);
        const int2 n2 = S.info[(size_t)min(s + 2, last) * TP_LANES + t];
        if (has && !own) {
            red4[t] = make_float4(p[0], p[1], p[2], p[3]);
            red2[t] = make_float2(p[4], p[5]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
        if (own) {
            for (int d0 = 1; d0 < nd; d0 += 4) {
                float4 q4[4];
                float2 q2[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t lu = t + (uint32_t)min(d0 + u, nd - 1);   // in-row lane (repeats masked below)
                    q4[u] = red4[lu];
                    q2[u] = red2[lu];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (d0 + u < nd) {
                        w[0] -= q4[u].x; w[1] -= q4[u].y; w[2] -= q4[u].z; w[3] -= q4[u].w;
                        w[4] -= q2[u].x; w[5] -= q2[u].y;
                    }
                }
            }
#pragma unroll
            for (int a = 0; a < 6; ++a) pw[a] = w[a];
            const uint32_t sl = (uint32_t)k & (TP_RING - 1);
            ring4[sl] = make_float4(w[0], w[1], w[2], w[3]);
            ring2[sl] = make_float2(w[4], w[5]);
        }
        pk = own ? k : -1;
        __syncthreads();
        cur = nxt;
        nxt = n2;
    }
    if (pk >= 0) row_st<r>(S.res, nr, (uint32_t)pk, pw);
}

__global__ void __launch_bounds__(TP_LANES) tp_sweep1_kernel(TaskDev S)
{
    extern __shared__ float4 tp1_lds[];
    if (blockIdx.x == 0) tp_sweep1_body<0>(S, tp1_lds);
    else tp_sweep1_body<1>(S, tp1_lds);
}

// Z_P = L_PP^-1 of every lower-triangular diagonal block (row-major)
__global__ void diag_inverse_kernel(int N, const float *Ldiag, float *Z)
{
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= N) return;
    float Lm[36], Zm[36];
    for (int q = 0; q < 36; ++q) { Lm[q] = Ldiag[(int64_t)P * 36 + q]; Zm[q] = 0.f; }
    for (int c = 0; c < 6; ++c) {
        Zm[6 * c + c] = 1.0f / Lm[6 * c + c];
        for (int r = c + 1; r < 6; ++r) {
            float s = 0.f;
            for (int q = c; q < r; ++q) s += Lm[6 * r + q] * Zm[6 * q + c];
            Zm[6 * r + c] = -s / Lm[6 * r + r];
        }
    }
    for (int q = 0; q < 36; ++q) Z[(int64_t)P * 36 + q] = Zm[q];
}

int k_diag_inverse(fsd_ctx *ctx)
{
    hipLaunchKernelGGL(diag_inverse_kernel, dim3(nblk(ctx->N, 128)), dim3(128), 0, ctx->stream, ctx->N, ctx->d_Ldiag,
                       ctx->d_tp_Z);
    FSD_HIP(hipGetLastError());
    return 0;
}

// Lane arrays of a sweep from its row placement (the rules of task_plan_lanes):
// a row's first lane holds (src of its first dependency, k | nd << TP_ND_SHIFT),
// its other lanes (src of dependency j, -1); bsrc = (factor block of the
// dependency, particle whose diagonal block scales it).  Lanes no row uses
// keep the -1 they were cleared to.  place = [lane: N][step: N][send: nsteps].
__global__ void tp_lanes_kernel(int N, const int *rowptr, const int *cols, const int *src, int forward, int R,
                                const int *seq, const int *place, int2 *info, int2 *bsrc)
{
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= N) return;
    const int g = place[P], S = place[N + P], r0 = rowptr[P], n = rowptr[P + 1] - r0;
    const int first = place[2 * N + S] - R;   // lowest sequence number still in the LDS ring at step S
    const int own = seq[P] | (n << TP_ND_SHIFT);
    if (n == 0) info[g] = make_int2(-1, own);
    for (int j = 0; j < n; ++j) {
        const int e = r0 + j, Q = cols[e], kq = seq[Q];
        const int s = kq >= first ? (kq & (R - 1)) : (TP_GLOBAL | kq);
        info[g + j] = make_int2(s, j == 0 ? own : -1);
        bsrc[g + j] = make_int2(src ? src[e] : e, forward ? P : Q);
    }
}

// lane blocks: M = Z_d L_e (forward) or its transpose (backward)
__global__ void tp_pack_kernel(int64_t lanes, const int2 *bsrc, const float *Lblk, const float *Z, float4 *blk,
                               int transpose)
{
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= lanes) return;
    const int2 sr = bsrc[g];
    if (sr.x < 0) return;
    const float *L = Lblk + (int64_t)sr.x * 36;
    const float *Zd = Z + (int64_t)sr.y * 36;
    float M[36];
    for (int a = 0; a < 6; ++a)
        for (int c = 0; c < 6; ++c) {
            float s = 0.f;
            for (int q = 0; q <= a; ++q) s += Zd[6 * a + q] * L[6 * q + c];
            if (transpose) M[6 * c + a] = s;
            else M[6 * a + c] = s;
        }
    for (int i = 0; i < 9; ++i)
        blk[tp_idx((uint32_t)g, i, 9)] = make_float4(M[4 * i], M[4 * i + 1], M[4 * i + 2], M[4 * i + 3]);
}

// forward inputs: inp[k_f(P)] = Z_P b_r[perm P]
__global__ void tp_pre_kernel(int N, const int *seq, const int *perm, const float *Z, const float *b0,
                              const float *b1, int nrhs, float4 *inp)
{
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= N) return;
    const float *Zp = Z + (int64_t)P * 36;
    const int64_t src = 6 * (int64_t)perm[P];
    float a[6], b[6];
    for (int r = 0; r < 6; ++r) {
        float sa = 0.f, sb = 0.f;
        for (int c = 0; c <= r; ++c) {
            sa += Zp[6 * r + c] * b0[src + c];
            if (nrhs == 2) sb += Zp[6 * r + c] * b1[src + c];
        }
        a[r] = sa;
        b[r] = sb;
    }
    const int k = seq[P];
    inp[k] = make_float4(a[0], a[1], a[2], a[3]);
    inp[N + k] = make_float4(a[4], a[5], b[0], b[1]);
    inp[2 * N + k] = make_float4(b[2], b[3], b[4], b[5]);
}

// backward inputs from forward results: inp_b[k_b(P)] = res_f[k_f(P)]
__global__ void tp_mid_kernel(int N, const int *seq_f, const int *seq_b, const float4 *res_f, float4 *inp_b)
{
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= N) return;
    const int kf = seq_f[P], kb = seq_b[P];
    for (int i = 0; i < 3; ++i) inp_b[i * N + kb] = res_f[i * N + kf];
}

// outputs: y_r[6 perm P + a] = (Z_P^T u_P)_a with u_P = res_b[k_b(P)]
__global__ void tp_post_kernel(int N, const int *seq, const int *perm, const float *Z, const float4 *res, int nrhs,
                               float *y0, float *y1)
{
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= N) return;
    const int k = seq[P];
    const float4 w0 = res[k], w1 = res[N + k];
    const float ua[6] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y};
    float ub[6] = {w1.z, w1.w, 0.f, 0.f, 0.f, 0.f};
    if (nrhs == 2) {
        const float4 w2 = res[2 * N + k];
        ub[2] = w2.x; ub[3] = w2.y; ub[4] = w2.z; ub[5] = w2.w;
    }
    const float *Zp = Z + (int64_t)P * 36;
    const int64_t dst = 6 * (int64_t)perm[P];
    for (int a = 0; a < 6; ++a) {
        float sa = 0.f, sb = 0.f;
        for (int c = a; c < 6; ++c) {
            sa += Zp[6 * c + a] * ua[c];
            sb += Zp[6 * c + a] * ub[c];
        }
        y0[dst + a] = sa;
        if (nrhs == 2) y1[dst + a] = sb;
    }
}

// Saddle preconditioner with one right-hand side.  Saddle_Preconditioner
// (Saddle.cu:171-264) forms G b_F and G b_U and combines them as
// x_F = b_F - G b_F + G b_U, x_U = G b_F - G b_U; G is linear, so with
// d = b_F - b_U this is x_F = b_F - G d, x_U = G d, and one sweep pair
// suffices.  Forward inputs inp[k_f(P)] = Z_P d[perm P] (right-hand side 0).
__global__ void tp_pre_diff_kernel(int N, const int *seq, const int *perm, const float *Z, const float *bF,
                                   const float *bU, float4 *inp)
{
    // six threads per particle: thread r forms component r (Z_P is lower triangular)
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 6 * (int64_t)N) return;
    const int P = (int)(q / 6), r = (int)(q - 6 * (int64_t)P);
    const float *Zr = Z + (int64_t)P * 36 + 6 * r;
    const int64_t src = 6 * (int64_t)perm[P];
    float sa = 0.f;
    for (int c = 0; c <= r; ++c) sa += Zr[c] * (bF[src + c] - bU[src + c]);
    const int k = seq[P];
    float *dst = r < 4 ? reinterpret_cast<float *>(inp + k) + r : reinterpret_cast<float *>(inp + N + k) + (r - 4);
    *dst = sa;
}

// backward inputs of right-hand side 0 from its forward results (two threads per particle:
// the float4 and the float2 of the row)
__global__ void tp_mid1_kernel(int N, const int *seq_f, const int *seq_b, const float4 *res_f, float4 *inp_b)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 2 * (int64_t)N) return;
    const int P = (int)(q >> 1);
    const int kf = seq_f[P], kb = seq_b[P];
    if ((q & 1) == 0) inp_b[kb] = res_f[kf];
    else reinterpret_cast<float2 *>(inp_b + N + kb)[0] = reinterpret_cast<const float2 *>(res_f + N + kf)[0];
}

// g = Z_P^T u_P (u: right-hand side 0 of the backward results), then
// x_F = b_F - g, x_S = b_S, x_U = g for particle perm[P]; six threads per
// particle, thread a forms component a (and copies stresslet component a < 5)
__global__ void tp_post_saddle_kernel(int N, const int *seq, const int *perm, const float *Z, const float4 *res,
                                      const float *b, float *x)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 6 * (int64_t)N) return;
    const int P = (int)(q / 6), a = (int)(q - 6 * (int64_t)P);
    const int k = seq[P];
    const float4 w0 = res[k];
    const float2 w1 = reinterpret_cast<const float2 *>(res + N + k)[0];
    const float u[6] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y};
    const float *Zp = Z + (int64_t)P * 36;
    const int64_t i = perm[P], n6 = 6 * (int64_t)N, n11 = 11 * (int64_t)N;
    float g = 0.f;
    for (int c = a; c < 6; ++c) g += Zp[6 * c + a] * u[c];
    x[6 * i + a] = b[6 * i + a] - g;
    x[n11 + 6 * i + a] = g;
    if (a < 5) x[n6 + 5 * i + a] = b[n6 + 5 * i + a];
}

// x = saddle preconditioner of b through the one-launch sweeps; returns 1 (nothing done)
// when the sweeps run level by level
int k_saddle_precond_tp(fsd_ctx *ctx, const float *b, float *x)
{
    if (!ctx->pch.tp_ok) return 1;
    const int N = ctx->N;
    hipEvent_t ev, es;
    prof_begin(ctx, PROF_TRSV, &ev);
    TaskDev F{ctx->pch.tp_f.nsteps, N, ctx->d_tp_finfo, ctx->d_tp_fblk, ctx->d_tp_finp, ctx->d_tp_fres};
    TaskDev B{ctx->pch.tp_b.nsteps, N, ctx->d_tp_binfo, ctx->d_tp_bblk, ctx->d_tp_binp, ctx->d_tp_bres};
    hipLaunchKernelGGL(tp_pre_diff_kernel, dim3(nblk(6 * (int64_t)N, 256)), dim3(256), 0, ctx->stream, N, ctx->d_tp_fseq,
                       ctx->d_perm, ctx->d_tp_Z, b, b + 11 * (int64_t)N, ctx->d_tp_finp);
    prof_begin(ctx, PROF_SWEEP, &es);
    hipLaunchKernelGGL(tp_sweep1_kernel, dim3(1), dim3(TP_LANES), TP1_LDS, ctx->stream, F);
    prof_end(ctx, PROF_SWEEP, es);
    hipLaunchKernelGGL(tp_mid1_kernel, dim3(nblk(2 * (int64_t)N, 256)), dim3(256), 0, ctx->stream, N, ctx->d_tp_fseq,
                       ctx->d_tp_bseq, ctx->d_tp_fres, ctx->d_tp_binp);
    prof_begin(ctx, PROF_SWEEP, &es);
    hipLaunchKernelGGL(tp_sweep1_kernel, dim3(1), dim3(TP_LANES), TP1_LDS, ctx->stream, B);
    prof_end(ctx, PROF_SWEEP, es);
    hipLaunchKernelGGL(tp_post_saddle_kernel, dim3(nblk(6 * (int64_t)N, 256)), dim3(256), 0, ctx->stream, N, ctx->d_tp_bseq,
                       ctx->d_perm, ctx->d_tp_Z, ctx->d_tp_bres, b, x);
    FSD_HIP(hipGetLastError());
    prof_end(ctx, PROF_TRSV, ev);
    return 0;
}

// whether the sweep kernel's LDS fits on this device (tp_R = 0: level by level)
int k_sweep_setup(fsd_ctx *ctx)
{
    int dev = 0, lds = 0;
    ctx->tp_R = 0;
    FSD_HIP(hipGetDevice(&dev));
    FSD_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const int bytes = (int)sizeof(float4) * 3 * (TP_RING + TP_LANES);
    if (lds < bytes) return 0;
    FSD_HIP(hipFuncSetAttribute((const void *)tp_sweep_kernel<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                bytes));
    FSD_HIP(hipFuncSetAttribute((const void *)tp_sweep1_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                TP1_LDS));
    ctx->tp_R = TP_RING;
    return 0;
}

// after IC(0): inverse diagonal blocks and the lane blocks of both sweeps
int k_sweep_pack(fsd_ctx *ctx)
{
    const int N = ctx->N;
    hipLaunchKernelGGL(diag_inverse_kernel, dim3(nblk(N, 128)), dim3(128), 0, ctx->stream, N, ctx->d_Ldiag,
                       ctx->d_tp_Z);
    const int64_t lf = (int64_t)ctx->pch.tp_f.nsteps * TP_LANES, lb = (int64_t)ctx->pch.tp_b.nsteps * TP_LANES;
    // lane arrays from the row placements
    FSD_HIP(hipMemsetAsync(ctx->d_tp_finfo, 0xFF, sizeof(int2) * lf, ctx->stream));
    FSD_HIP(hipMemsetAsync(ctx->d_tp_fbsrc, 0xFF, sizeof(int2) * lf, ctx->stream));
    FSD_HIP(hipMemsetAsync(ctx->d_tp_binfo, 0xFF, sizeof(int2) * lb, ctx->stream));
    FSD_HIP(hipMemsetAsync(ctx->d_tp_bbsrc, 0xFF, sizeof(int2) * lb, ctx->stream));
    hipLaunchKernelGGL(tp_lanes_kernel, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, N, ctx->d_lrow, ctx->d_lcol,
                       (const int *)nullptr, 1, ctx->tp_R, ctx->d_tp_fseq, ctx->d_tp_fplace, ctx->d_tp_finfo,
                       ctx->d_tp_fbsrc);
    hipLaunchKernelGGL(tp_lanes_kernel, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, N, ctx->d_urow, ctx->d_ucol,
                       ctx->d_upos, 0, ctx->tp_R, ctx->d_tp_bseq, ctx->d_tp_bplace, ctx->d_tp_binfo, ctx->d_tp_bbsrc);
    if (lf)
        hipLaunchKernelGGL(tp_pack_kernel, dim3(nblk(lf, 256)), dim3(256), 0, ctx->stream, lf, ctx->d_tp_fbsrc,
                           ctx->d_Lblk, ctx->d_tp_Z, ctx->d_tp_fblk, 0);
    if (lb)
        hipLaunchKernelGGL(tp_pack_kernel, dim3(nblk(lb, 256)), dim3(256), 0, ctx->stream, lb, ctx->d_tp_bbsrc,
                           ctx->d_Lblk, ctx->d_tp_Z, ctx->d_tp_bblk, 1);
    FSD_HIP(hipGetLastError());
    return 0;
}

// y_r = G b_r for r < n_rhs (b1/y1 used when n_rhs == 2)
int k_precond_apply2(fsd_ctx *ctx, const float *b0, const float *b1, float *y0, float *y1, int n_rhs)
{
    const int N = ctx->N;
    hipEvent_t ev;
    prof_begin(ctx, PROF_TRSV, &ev);
    if (ctx->pch.tp_ok) {
        const size_t lds = sizeof(float4) * 3 * (TP_RING + TP_LANES);
        TaskDev F{ctx->pch.tp_f.nsteps, N, ctx->d_tp_finfo, ctx->d_tp_fblk, ctx->d_tp_finp, ctx->d_tp_fres};
        TaskDev B{ctx->pch.tp_b.nsteps, N, ctx->d_tp_binfo, ctx->d_tp_bblk, ctx->d_tp_binp, ctx->d_tp_bres};
        hipLaunchKernelGGL(tp_pre_kernel, dim3(nblk(N, 128)), dim3(128), 0, ctx->stream, N, ctx->d_tp_fseq,
                           ctx->d_perm, ctx->d_tp_Z, b0, b1, n_rhs, ctx->d_tp_finp);
        hipEvent_t es;
        prof_begin(ctx, PROF_SWEEP, &es);
        if (n_rhs == 2 && ctx->tp_joint)
            hipLaunchKernelGGL(tp_sweep_kernel<2>, dim3(1), dim3(TP_LANES), lds, ctx->stream, F);
        else
            hipLaunchKernelGGL(tp_sweep1_kernel, dim3(n_rhs), dim3(TP_LANES), TP1_LDS, ctx->stream, F);
        prof_end(ctx, PROF_SWEEP, es);
        hipLaunchKernelGGL(tp_mid_kernel, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, N, ctx->d_tp_fseq,
                           ctx->d_tp_bseq, ctx->d_tp_fres, ctx->d_tp_binp);
        prof_begin(ctx, PROF_SWEEP, &es);
        if (n_rhs == 2 && ctx->tp_joint)
            hipLaunchKernelGGL(tp_sweep_kernel<2>, dim3(1), dim3(TP_LANES), lds, ctx->stream, B);
        else
            hipLaunchKernelGGL(tp_sweep1_kernel, dim3(n_rhs), dim3(TP_LANES), TP1_LDS, ctx->stream, B);
        prof_end(ctx, PROF_SWEEP, es);
        hipLaunchKernelGGL(tp_post_kernel, dim3(nblk(N, 128)), dim3(128), 0, ctx->stream, N, ctx->d_tp_bseq,
                           ctx->d_perm, ctx->d_tp_Z, ctx->d_tp_bres, n_rhs, y0, y1);
        FSD_HIP(hipGetLastError());
        prof_end(ctx, PROF_TRSV, ev);
        return 0;
    }
    for (int l = 0; l < ctx->pch.nlev_f; ++l) {
        const int b = ctx->pch.lev_f_ptr[l], n = ctx->pch.lev_f_ptr[l + 1] - b;
        hipLaunchKernelGGL(trsv_fwd_level_kernel, dim3(n), dim3(64), 0, ctx->stream, ctx->d_lev_f + b, N, n_rhs,
                           ctx->d_lrow, ctx->d_lcol, ctx->d_Lblk, ctx->d_Ldiag, ctx->d_perm, b0, b1, ctx->d_tmp6a);
    }
    for (int l = 0; l < ctx->pch.nlev_b; ++l) {
        const int b = ctx->pch.lev_b_ptr[l], n = ctx->pch.lev_b_ptr[l + 1] - b;
        hipLaunchKernelGGL(trsv_bwd_level_kernel, dim3(n), dim3(64), 0, ctx->stream, ctx->d_lev_b + b, N, n_rhs,
                           ctx->d_urow, ctx->d_ucol, ctx->d_upos, ctx->d_Lblk, ctx->d_Ldiag, ctx->d_perm,
                           ctx->d_tmp6a, ctx->d_tmp6b, y0, y1);
    }
    FSD_HIP(hipGetLastError());
    prof_end(ctx, PROF_TRSV, ev);
    return 0;
}

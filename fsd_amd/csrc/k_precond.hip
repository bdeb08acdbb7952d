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

// ------------------------------------- single-workgroup sweeps (one launch)
// Both sweeps of one application run in one workgroup of SW_THREADS threads;
// __syncthreads orders the levels, so no workgroup ever waits on another.
// Three lanes own one row (lane c: output components 2c and 2c+1) for both
// right-hand sides; a wave holds 21 rows (lane 63 idle).  Inputs, blocks,
// inverse diagonal blocks and dependency positions are stored in sweep order
// (jagged-diagonal inside a level) and results are written in sweep order:
// every load of a row is issued at once from its position, and the only
// dependent loads are the dependencies' values, one batch per level.
struct SweepDev {
    int nlev, levrow, levjd, rec;   // offsets (ints) into the schedule table
    const int *col;                 // entry -> position of the dependency
    const float *blk;               // entry -> 6x6 block, row a feeds output component a
    const float *dinv;              // position -> inverse diagonal block (row-major)
    const float *in;                // 2 x 6N input in sweep order
    float *res;                     // 2 x 6N results in sweep order
    const int *scatter;             // position -> 6-vector index of the copy
    float *copy0, *copy1;           // copies of the results (right-hand sides 0 and 1)
};

#define SW_JC 8   // dependency positions loaded per batch
#define SW_JB 4   // dependency values loaded per batch

// 32-bit byte offsets from a uniform base: global loads in saddr form (the
// host checks that every offset fits in 32 bits)
template <typename T>
__device__ __forceinline__ T ld_off(const void *base, uint32_t byte_off)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte_off);
}
template <typename T>
__device__ __forceinline__ void st_off(void *base, uint32_t byte_off, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + byte_off) = v;
}

template <int NRHS>
__device__ __forceinline__ void sweep_levels(const int *s_tab, const SweepDev &S, uint32_t N6b, int slot, int nslot,
                                             int c, int gbase, bool lane_ok)
{
    for (int l = 0; l < S.nlev; ++l) {
        const int r0 = s_tab[S.levrow + l], r1 = s_tab[S.levrow + l + 1];
        const int j0 = s_tab[S.levjd + l], nj = s_tab[S.levjd + l + 1] - j0;
        for (int base = r0; base < r1; base += nslot) {
            const int k = base + slot;
            const bool act = lane_ok && k < r1;
            const int kl = k - r0;
            float2 v0 = make_float2(0.f, 0.f), v1 = make_float2(0.f, 0.f);
            float4 D0 = make_float4(0.f, 0.f, 0.f, 0.f), D1 = D0, D2 = D0;
            const uint32_t kofs = 24u * (uint32_t)k + 8u * (uint32_t)c;   // bytes of components 2c, 2c+1 of row k
            if (act) {
                v0 = ld_off<float2>(S.in, kofs);
                if (NRHS == 2) v1 = ld_off<float2>(S.in, N6b + kofs);
                const uint32_t dofs = 144u * (uint32_t)k + 48u * (uint32_t)c;
                D0 = ld_off<float4>(S.dinv, dofs);
                D1 = ld_off<float4>(S.dinv, dofs + 16u);
                D2 = ld_off<float4>(S.dinv, dofs + 32u);
                for (int jb = 0; jb < nj; jb += SW_JC) {
                    // dependency positions of up to SW_JC blocks at once, then their values SW_JB at a time
                    uint32_t eo[SW_JC], qo[SW_JC];
                    int n = 0;
#pragma unroll
                    for (int t = 0; t < SW_JC; ++t) {
                        const int j = j0 + jb + t;
                        const bool has = (jb + t < nj) && kl < s_tab[S.rec + 2 * j + 1];
                        n += has;
                        const uint32_t e = has ? (uint32_t)(s_tab[S.rec + 2 * j] + kl) : 0u;  // entry 0 exists
                        eo[t] = 144u * e + 48u * (uint32_t)c;
                        qo[t] = 24u * (uint32_t)ld_off<int>(S.col, 4u * e);
                    }
#pragma unroll
                    for (int tb = 0; tb < SW_JC; tb += SW_JB) {
                        if (tb >= n) break;
#pragma unroll
                        for (int t = tb; t < tb + SW_JB; ++t) {
                            // row 2c: b0, b1.xy; row 2c+1: b1.zw, b2
                            const float4 b0 = ld_off<float4>(S.blk, eo[t]);
                            const float4 b1 = ld_off<float4>(S.blk, eo[t] + 16u);
                            const float4 b2 = ld_off<float4>(S.blk, eo[t] + 32u);
                            // slots t >= n read a valid row whose value may be stale: selected away below
                            const float2 x01 = ld_off<float2>(S.res, qo[t]);
                            const float2 x23 = ld_off<float2>(S.res, qo[t] + 8u);
                            const float2 x45 = ld_off<float2>(S.res, qo[t] + 16u);
                            const float p0 = b0.x * x01.x + b0.y * x01.y + b0.z * x23.x + b0.w * x23.y +
                                             b1.x * x45.x + b1.y * x45.y;
                            const float p1 = b1.z * x01.x + b1.w * x01.y + b2.x * x23.x + b2.y * x23.y +
                                             b2.z * x45.x + b2.w * x45.y;
                            v0.x -= (t < n) ? p0 : 0.f;
                            v0.y -= (t < n) ? p1 : 0.f;
                            if (NRHS == 2) {
                                const float2 y01 = ld_off<float2>(S.res, N6b + qo[t]);
                                const float2 y23 = ld_off<float2>(S.res, N6b + qo[t] + 8u);
                                const float2 y45 = ld_off<float2>(S.res, N6b + qo[t] + 16u);
                                const float p2 = b0.x * y01.x + b0.y * y01.y + b0.z * y23.x + b0.w * y23.y +
                                                 b1.x * y45.x + b1.y * y45.y;
                                const float p3 = b1.z * y01.x + b1.w * y01.y + b2.x * y23.x + b2.y * y23.y +
                                                 b2.z * y45.x + b2.w * y45.y;
                                v1.x -= (t < n) ? p2 : 0.f;
                                v1.y -= (t < n) ? p3 : 0.f;
                            }
                        }
                    }
                }
            }
            // the row's six components from its three lanes
            float u0[6], u1[6];
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                u0[2 * m] = __shfl(v0.x, gbase + m);
                u0[2 * m + 1] = __shfl(v0.y, gbase + m);
                if (NRHS == 2) {
                    u1[2 * m] = __shfl(v1.x, gbase + m);
                    u1[2 * m + 1] = __shfl(v1.y, gbase + m);
                }
            }
            if (act) {
                // D0, D1.xy: row 2c of the inverse; D1.zw, D2: row 2c+1
                float2 w0, w1 = make_float2(0.f, 0.f);
                w0.x = D0.x * u0[0] + D0.y * u0[1] + D0.z * u0[2] + D0.w * u0[3] + D1.x * u0[4] + D1.y * u0[5];
                w0.y = D1.z * u0[0] + D1.w * u0[1] + D2.x * u0[2] + D2.y * u0[3] + D2.z * u0[4] + D2.w * u0[5];
                if (NRHS == 2) {
                    w1.x = D0.x * u1[0] + D0.y * u1[1] + D0.z * u1[2] + D0.w * u1[3] + D1.x * u1[4] + D1.y * u1[5];
                    w1.y = D1.z * u1[0] + D1.w * u1[1] + D2.x * u1[2] + D2.y * u1[3] + D2.z * u1[4] + D2.w * u1[5];
                }
                st_off<float2>(S.res, kofs, w0);
                if (NRHS == 2) st_off<float2>(S.res, N6b + kofs, w1);
                const uint32_t o = 24u * (uint32_t)ld_off<int>(S.scatter, 4u * (uint32_t)k) + 8u * (uint32_t)c;
                st_off<float2>(S.copy0, o, w0);
                if (NRHS == 2) st_off<float2>(S.copy1, o, w1);
            }
        }
        __syncthreads();
    }
}

template <int NRHS>
__global__ void __launch_bounds__(SW_THREADS) trsv_sweeps_kernel(const int *tab, int ntab, SweepDev F, SweepDev B,
                                                                 uint32_t N6b)
{
    extern __shared__ int s_tab[];
    for (int t = threadIdx.x; t < ntab; t += blockDim.x) s_tab[t] = tab[t];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int g = lane / 3;
    const bool lane_ok = lane < 63;
    const int slot = (threadIdx.x >> 6) * 21 + g;
    const int nslot = (blockDim.x >> 6) * 21;
    const int c = lane - 3 * g;
    // forward: y = L^-1 P b in forward order, copied to backward order
    sweep_levels<NRHS>(s_tab, F, N6b, slot, nslot, c, 3 * g, lane_ok);
    // backward: x = L^-T y in backward order, copied to the outputs in particle order
    sweep_levels<NRHS>(s_tab, B, N6b, slot, nslot, c, 3 * g, lane_ok);
}

// inputs of the forward sweep: in[r][6k + a] = b_r[6 fin[k] + a]
__global__ void sweep_gather_kernel(int N, const int *fin, const float *b0, const float *b1, int nrhs, float *in)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t N6 = 6 * (int64_t)N;
    if (t >= N6) return;
    const int64_t k = t / 6;
    const int a = (int)(t - 6 * k);
    const int64_t s = 6 * (int64_t)fin[k] + a;
    in[t] = b0[s];
    if (nrhs == 2) in[N6 + t] = b1[s];
}

// blocks in sweep order: forward copies L_PQ, backward stores (L_QP)^T
__global__ void sweep_pack_blocks_kernel(int64_t n, const int *src, const float *Lblk, float *dst, int transpose)
{
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= n * 36) return;
    const int64_t e = idx / 36;
    const int q = (int)(idx - e * 36);
    const int a = q / 6, c = q - 6 * a;
    const int64_t s = (int64_t)src[e] * 36;
    dst[idx] = transpose ? Lblk[s + 6 * c + a] : Lblk[s + q];
}

// inverse Z = L_PP^-1 of every lower-triangular diagonal block: forward order
// stores Z, backward order stores Z^T
__global__ void sweep_pack_dinv_kernel(int N, const int *f_order, const int *b_order, const float *Ldiag, float *fd,
                                       float *bd)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * N) return;
    const bool fwd = t < N;
    const int k = fwd ? t : t - N;
    const int P = fwd ? f_order[k] : b_order[k];
    float Lm[36], Z[36];
    for (int q = 0; q < 36; ++q) { Lm[q] = Ldiag[(int64_t)P * 36 + q]; Z[q] = 0.f; }
    for (int c = 0; c < 6; ++c) {
        Z[6 * c + c] = 1.0f / Lm[6 * c + c];
        for (int r = c + 1; r < 6; ++r) {
            float s = 0.f;
            for (int q = c; q < r; ++q) s += Lm[6 * r + q] * Z[6 * q + c];
            Z[6 * r + c] = -s / Lm[6 * r + r];
        }
    }
    float *dst = (fwd ? fd : bd) + (int64_t)k * 36;
    for (int a = 0; a < 6; ++a)
        for (int c = 0; c < 6; ++c) dst[6 * a + c] = fwd ? Z[6 * a + c] : Z[6 * c + a];
}

int k_sweep_pack(fsd_ctx *ctx)
{
    const int N = ctx->N;
    const int64_t n = ctx->nlow;
    if (n) {
        hipLaunchKernelGGL(sweep_pack_blocks_kernel, dim3(nblk(n * 36, 256)), dim3(256), 0, ctx->stream, n,
                           ctx->d_sw_fsrc, ctx->d_Lblk, ctx->d_sw_fblk, 0);
        hipLaunchKernelGGL(sweep_pack_blocks_kernel, dim3(nblk(n * 36, 256)), dim3(256), 0, ctx->stream, n,
                           ctx->d_sw_bsrc, ctx->d_Lblk, ctx->d_sw_bblk, 1);
    }
    hipLaunchKernelGGL(sweep_pack_dinv_kernel, dim3(nblk(2 * (int64_t)N, 128)), dim3(128), 0, ctx->stream, N,
                       ctx->d_sw_forder, ctx->d_sw_border, ctx->d_Ldiag, ctx->d_sw_fdinv, ctx->d_sw_bdinv);
    FSD_HIP(hipGetLastError());
    return 0;
}

// y_r = G b_r for r < n_rhs (b1/y1 used when n_rhs == 2)
int k_precond_apply2(fsd_ctx *ctx, const float *b0, const float *b1, float *y0, float *y1, int n_rhs)
{
    const int N = ctx->N;
    hipEvent_t ev;
    prof_begin(ctx, PROF_TRSV, &ev);
    if (ctx->sw_ok) {
        const int64_t N6 = 6 * (int64_t)N;
        float *bin = ctx->d_sw_work, *yf = bin + 2 * N6, *yb = yf + 2 * N6, *x = yb + 2 * N6;
        hipLaunchKernelGGL(sweep_gather_kernel, dim3(nblk(N6, 256)), dim3(256), 0, ctx->stream, N, ctx->d_sw_fin, b0,
                           b1, n_rhs, bin);
        SweepDev F{ctx->sw_f.nlev, ctx->sw_f.levrow, ctx->sw_f.levjd, ctx->sw_f.rec, ctx->d_sw_fcol, ctx->d_sw_fblk,
                   ctx->d_sw_fdinv, bin, yf, ctx->d_sw_fyb, yb, yb + N6};
        SweepDev B{ctx->sw_b.nlev, ctx->sw_b.levrow, ctx->sw_b.levjd, ctx->sw_b.rec, ctx->d_sw_bcol, ctx->d_sw_bblk,
                   ctx->d_sw_bdinv, yb, x, ctx->d_sw_bout, y0, (n_rhs == 2) ? y1 : y0};
        const size_t lds = sizeof(int) * (size_t)ctx->sw_ntab;
        if (n_rhs == 2)
            hipLaunchKernelGGL(trsv_sweeps_kernel<2>, dim3(1), dim3(SW_THREADS), lds, ctx->stream, ctx->d_sw_tab,
                               ctx->sw_ntab, F, B, (uint32_t)(4 * N6));
        else
            hipLaunchKernelGGL(trsv_sweeps_kernel<1>, dim3(1), dim3(SW_THREADS), lds, ctx->stream, ctx->d_sw_tab,
                               ctx->sw_ntab, F, B, (uint32_t)(4 * N6));
        FSD_HIP(hipGetLastError());
        prof_end(ctx, PROF_TRSV, ev);
        return 0;
    }
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

// Preconditioner application by Jacobi iterations on the triangular factors.
//
// G = P^T L^-T L^-1 P is applied without the level-by-level dependency chain:
// with Z_P = L_PP^-1 and the folded blocks F_PQ = Z_P L_PQ (Q < P), the
// forward solve L y = t is y = c - F y with c = Z t, and the backward solve
// L^T x = y is x = Z^T u with u = y - F^T u.  Both are iterated k times from
// y = c and u = y (Anzt, Chow & Dongarra, "Iterative sparse triangular solves
// for preconditioning", Euro-Par 2015).  Every iteration reads each block of
// the factor once and is fully parallel over the N block rows, so one
// application costs 2k + 2 short launches over all CUs instead of two chains
// of ~120 dependent steps on one CU.  For a fixed k the result is a fixed
// linear operator (a truncated Neumann series of the exact solves), so GMRES
// needs no flexible variant; k -> depth of the DAG gives the exact solves.
//
// Layout: every per-row vector is [N][6] in permuted order P; the blocks F are
// [lower block][6][6] row-major in the order of the factor's lower pattern.
// A 64-lane wave owns 10 rows, six lanes per row (lane a computes component a),
// so the six lanes of a row read one block's 144 contiguous bytes.
#include "fsd_internal.h"

namespace {
constexpr int JROWS = 10;   // block rows per 64-lane wave (lanes 60..63 idle)

__device__ __forceinline__ bool jac_row(int N, int &P, int &a)
{
    const int lane = threadIdx.x & 63;
    // rows in one contiguous range per XCD (cache lines shared by neighbouring rows stay in one L2)
    const int wave = (int)((xcd_remap(blockIdx.x, gridDim.x) * (int64_t)blockDim.x + threadIdx.x) >> 6);
    const int g = lane / 6;
    a = lane - 6 * g;
    P = wave * JROWS + g;
    return g < JROWS && P < N;
}

inline int jac_blocks(int N) { return nblk(nblk(N, JROWS) * 64, 256); }

__device__ __forceinline__ float dot6(const float *f, const float *y)
{
    const float2 f0 = *(const float2 *)f, f1 = *(const float2 *)(f + 2), f2 = *(const float2 *)(f + 4);
    const float2 y0 = *(const float2 *)y, y1 = *(const float2 *)(y + 2), y2 = *(const float2 *)(y + 4);
    return f0.x * y0.x + f0.y * y0.y + f1.x * y1.x + f1.y * y1.y + f2.x * y2.x + f2.y * y2.y;
}
}  // namespace

// row index of every lower block (the block rows are the CSR rows lrow)
__global__ void jac_rows_kernel(int N, const int *lrow, int *rowidx)
{
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= N) return;
    for (int e = lrow[P]; e < lrow[P + 1]; ++e) rowidx[e] = P;
}

// F_e = Z_P L_e for the lower block e of row P; six threads per block, thread a
// forms row a (Z lower triangular: row a of Z has entries 0..a)
__global__ void jac_fold_kernel(int64_t nlow, const int *rowidx, const float *Z, const float *Lblk, float *F)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= 6 * nlow) return;
    const int64_t e = t / 6;
    const int a = (int)(t - 6 * e);
    const float *Zp = Z + (int64_t)rowidx[e] * 36 + 6 * a;
    const float *L = Lblk + e * 36;
    float r[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int q = 0; q <= a; ++q) {
        const float z = Zp[q];
        for (int c = 0; c < 6; ++c) r[c] += z * L[6 * q + c];
    }
    float *o = F + e * 36 + 6 * a;
    for (int c = 0; c < 6; ++c) o[c] = r[c];
}

// c_P = Z_P t_P with t_P = (b0 - b1)[perm P] (b1 may be null); y0 = c as well
__global__ void jac_pre_kernel(int N, const int *perm, const float *Z, const float *b0, const float *b1, float *c)
{
    int P, a;
    if (!jac_row(N, P, a)) return;
    const int64_t i = perm[P];
    const float *Zp = Z + (int64_t)P * 36 + 6 * a;
    float v = 0.f;
    for (int q = 0; q <= a; ++q) {
        float t = b0[6 * i + q];
        if (b1) t -= b1[6 * i + q];
        v += Zp[q] * t;
    }
    c[6 * (int64_t)P + a] = v;
}

// forward iteration: yout_P = c_P - sum_{Q < P} F_PQ yin_Q
__global__ void __launch_bounds__(256) jac_fwd_kernel(int N, const int *lrow, const int *lcol, const float *F,
                                                      const float *c, const float *yin, float *yout)
{
    int P, a;
    if (!jac_row(N, P, a)) return;
    float v = c[6 * (int64_t)P + a];
    const int e1 = lrow[P + 1];
    for (int e = lrow[P]; e < e1; ++e) v -= dot6(F + (int64_t)e * 36 + 6 * a, yin + 6 * (int64_t)lcol[e]);
    yout[6 * (int64_t)P + a] = v;
}

// backward iteration: uout_P = y_P - sum_{Q > P} F_QP^T uin_Q (column a of block (Q, P))
__global__ void __launch_bounds__(256) jac_bwd_kernel(int N, const int *urow, const int *ucol, const int *upos,
                                                      const float *F, const float *y, const float *uin, float *uout)
{
    int P, a;
    if (!jac_row(N, P, a)) return;
    float v = y[6 * (int64_t)P + a];
    const int e1 = urow[P + 1];
    for (int e = urow[P]; e < e1; ++e) {
        const float *f = F + (int64_t)upos[e] * 36 + a;
        const float *u = uin + 6 * (int64_t)ucol[e];
        v -= f[0] * u[0] + f[6] * u[1] + f[12] * u[2] + f[18] * u[3] + f[24] * u[4] + f[30] * u[5];
    }
    uout[6 * (int64_t)P + a] = v;
}

// g = P^T Z^T u; mode 0: out = g (6N); mode 1: the saddle preconditioner's
// combination x_F = b_F - g, x_S = b_S, x_U = g (Saddle.cu:171-264 with
// g = G (b_F - b_U))
__global__ void jac_post_kernel(int N, const int *perm, const float *Z, const float *u, const float *b, float *x,
                                int mode)
{
    int P, a;
    if (!jac_row(N, P, a)) return;
    const float *Zp = Z + (int64_t)P * 36;
    const float *up = u + 6 * (int64_t)P;
    float g = 0.f;
    for (int c = a; c < 6; ++c) g += Zp[6 * c + a] * up[c];
    const int64_t i = perm[P], n6 = 6 * (int64_t)N, n11 = 11 * (int64_t)N;
    if (mode == 0) {
        x[6 * i + a] = g;
        return;
    }
    x[6 * i + a] = b[6 * i + a] - g;
    x[n11 + 6 * i + a] = g;
    if (a < 5) x[n6 + 5 * i + a] = b[n6 + 5 * i + a];
}

// Z = diagonal block inverses and F = Z L of the current factor
int k_jacobi_pack(fsd_ctx *ctx)
{
    const int N = ctx->N;
    const int64_t nlow = ctx->pch.nlow;
    {
        const int r = k_diag_inverse(ctx);
        if (r) return r;
    }
    if (nlow) {
        hipLaunchKernelGGL(jac_rows_kernel, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, N, ctx->d_lrow,
                           ctx->d_lrowidx);
        hipLaunchKernelGGL(jac_fold_kernel, dim3(nblk(6 * nlow, 256)), dim3(256), 0, ctx->stream, nlow,
                           ctx->d_lrowidx, ctx->d_tp_Z, ctx->d_Lblk, ctx->d_Fblk);
    }
    FSD_HIP(hipGetLastError());
    return 0;
}

// k forward and k backward iterations from c (already in buf[0]); returns the
// buffer that holds u
static const float *jac_sweeps(fsd_ctx *ctx, float *const buf[4])
{
    const int N = ctx->N, k = ctx->ic_jac_k;
    const dim3 g(jac_blocks(N)), b(256);
    const float *c = buf[0];
    const float *y = c;
    float *fo[2] = {buf[1], buf[2]};
    hipEvent_t e1;
    prof_begin(ctx, PROF_JAC, &e1);   // the k forward iterations, one event pair
    for (int it = 0; it < k; ++it) {
        float *out = fo[it & 1];
        hipLaunchKernelGGL(jac_fwd_kernel, g, b, 0, ctx->stream, N, ctx->d_lrow, ctx->d_lcol, ctx->d_Fblk, c, y, out);
        y = out;
    }
    prof_end_n(ctx, PROF_JAC, e1, k);
    // the backward iterates never overwrite y: they alternate between buf[3] and the forward buffer y does not use
    float *bo[2] = {buf[3], (y == buf[1]) ? buf[2] : buf[1]};
    const float *u = y;
    for (int it = 0; it < k; ++it) {
        float *out = bo[it & 1];
        hipLaunchKernelGGL(jac_bwd_kernel, g, b, 0, ctx->stream, N, ctx->d_urow, ctx->d_ucol, ctx->d_upos,
                           ctx->d_Fblk, y, u, out);
        u = out;
    }
    return u;
}

static int jac_apply(fsd_ctx *ctx, const float *b0, const float *b1, const float *bsaddle, float *x, int mode)
{
    const int N = ctx->N;
    hipEvent_t ev;
    prof_begin(ctx, PROF_TRSV, &ev);
    float *const buf[4] = {ctx->d_jac, ctx->d_jac + 6 * (int64_t)N, ctx->d_jac + 12 * (int64_t)N,
                           ctx->d_jac + 18 * (int64_t)N};
    hipLaunchKernelGGL(jac_pre_kernel, dim3(jac_blocks(N)), dim3(256), 0, ctx->stream, N, ctx->d_perm, ctx->d_tp_Z,
                       b0, b1, buf[0]);
    const float *u = jac_sweeps(ctx, buf);
    hipLaunchKernelGGL(jac_post_kernel, dim3(jac_blocks(N)), dim3(256), 0, ctx->stream, N, ctx->d_perm, ctx->d_tp_Z,
                       u, bsaddle, x, mode);
    FSD_HIP(hipGetLastError());
    prof_end(ctx, PROF_TRSV, ev);
    return 0;
}

// x = saddle preconditioner of b (17N) with g = G (b_F - b_U)
int k_saddle_precond_jacobi(fsd_ctx *ctx, const float *b, float *x)
{
    return jac_apply(ctx, b, b + 11 * (int64_t)ctx->N, b, x, 1);
}

// y = G b (6N)
int k_precond_apply_jacobi(fsd_ctx *ctx, const float *b, float *y)
{
    return jac_apply(ctx, b, nullptr, nullptr, y, 0);
}

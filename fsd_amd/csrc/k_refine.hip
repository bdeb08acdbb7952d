// FP64-residual iterative refinement of the saddle solve.
//
// The solution is accumulated in FP64.  Its residual r = b - A x is formed
// with the lubrication product x_F - R_FU x_U evaluated in FP64 on the FP32
// pair coefficients (the product whose near-contact terms cancel: R_FU's
// entries reach ~250 while a touching pair's relative velocity is small) and
// with the subtraction from b in FP64; the mobility rows M x_top + x_U come
// from the FP32 kernels (bounded entries, no cancellation).  Each correction
// A d = r is solved by the FP32 GMRES and added to x in FP64.  This wraps the
// reference's solve (Solvers.cu:55-126), which has no refinement.
#include "fsd_internal.h"

// R_FU in FP64 arithmetic, one thread per particle: out64 = xF64 - R U64
__global__ void rfu_apply64_kernel(int N, const double *U, const int *lub_n, const int *lub_j, const float4 *lub_a,
                                   const float4 *lub_b, const float *lub_self, int maxlub, const double *xF,
                                   double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const int n = lub_n[i];
    for (int k = 0; k < n; ++k) {
        const int64_t o = (int64_t)i * maxlub + k;
        const int j = lub_j[o];
        const float4 a = lub_a[o], bb = lub_b[o];
        const double *u = U + 6 * (int64_t)j;
        const double ux = u[0], uy = u[1], uz = u[2], wx = u[3], wy = u[4], wz = u[5];
        const double rx = a.x, ry = a.y, rz = a.z;
        const double dA = a.w, YA = bb.x, YB = bb.y, dC = bb.z, YC = bb.w;
        const double rdu = rx * ux + ry * uy + rz * uz;
        const double rdw = rx * wx + ry * wy + rz * wz;
        const double eux = rz * uy - ry * uz, euy = -rz * ux + rx * uz, euz = ry * ux - rx * uy;
        const double ewx = rz * wy - ry * wz, ewy = -rz * wx + rx * wz, ewz = ry * wx - rx * wy;
        acc[0] += dA * rdu * rx + YA * ux + YB * ewx;
        acc[1] += dA * rdu * ry + YA * uy + YB * ewy;
        acc[2] += dA * rdu * rz + YA * uz + YB * ewz;
        acc[3] += YB * eux + dC * rdw * rx + YC * wx;
        acc[4] += YB * euy + dC * rdw * ry + YC * wy;
        acc[5] += YB * euz + dC * rdw * rz + YC * wz;
    }
    const float *S = lub_self + (int64_t)i * 36;
    const double *ui = U + 6 * (int64_t)i;
    for (int r = 0; r < 6; ++r) {
        double s = acc[r];
        for (int c = 0; c < 6; ++c) s += (double)S[r * 6 + c] * ui[c];
        out[6 * (int64_t)i + r] = xF[6 * (int64_t)i + r] - s;
    }
}

// r = b - A x: rows [0, 11N) from the FP32 product y (M x_top + x_U), rows [11N, 17N) from
// the FP64 lubrication row ybot; also r32 = r (the next correction's right-hand side)
__global__ void resid64_kernel(int64_t n11, int64_t n, const float *b, const float *y, const double *ybot,
                               double *r, float *r32)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double v = (double)b[q] - (q < n11 ? (double)y[q] : ybot[q - n11]);
    r[q] = v;
    r32[q] = (float)v;
}

__global__ void to32_kernel(int64_t n, const double *x, float *y)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q < n) y[q] = (float)x[q];
}

__global__ void to64_kernel(int64_t n, const float *x, double *y)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q < n) y[q] = (double)x[q];
}

__global__ void add64_kernel(int64_t n, const float *d, double *x)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q < n) x[q] += (double)d[q];
}

// sums of squares of r (double) and b (float) into acc[0], acc[1] (zeroed by the caller)
__global__ void nrm64_kernel(int64_t n, const double *r, const float *b, double *acc)
{
    __shared__ double red[2][256];
    double s0 = 0.0, s1 = 0.0;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        s0 += r[q] * r[q];
        const double bv = b[q];
        s1 += bv * bv;
    }
    red[0][threadIdx.x] = s0;
    red[1][threadIdx.x] = s1;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(acc, red[0][0]);
        atomicAdd(acc + 1, red[1][0]);
    }
}

__global__ void zero64_kernel(double *x, int n)
{
    if ((int)threadIdx.x < n) x[threadIdx.x] = 0.0;
}

// r = b - A x64 with x32 = fl32(x64) for the FP32 mobility rows; |r|^2 and |b|^2 into norms
int k_residual64(fsd_ctx *ctx, const float4 *pos, const float *b, const double *x64, float *x32, float *y32,
                 double *ybot, double *r64, float *r32, double *norms)
{
    const int64_t N = ctx->N, n = 17 * N, n11 = 11 * N;
    hipLaunchKernelGGL(to32_kernel, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, n, x64, x32);
    {
        const int rc = k_mobility(ctx, pos, x32, y32, 3, x32 + n11);
        if (rc) return rc;
    }
    if (!ctx->lub_valid) { ctx->err = "lubrication pairs not built"; return FSD_E_STATE; }
    hipLaunchKernelGGL(rfu_apply64_kernel, dim3(nblk(N, 128)), dim3(128), 0, ctx->stream, (int)N, x64 + n11,
                       ctx->d_lub_n, ctx->d_lub_j, ctx->d_lub_a, ctx->d_lub_b, ctx->d_lub_self, ctx->maxlub, x64,
                       ybot);
    hipLaunchKernelGGL(resid64_kernel, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, n11, n, b, y32, ybot, r64, r32);
    hipLaunchKernelGGL(zero64_kernel, dim3(1), dim3(64), 0, ctx->stream, norms, 2);
    hipLaunchKernelGGL(nrm64_kernel, dim3(256), dim3(256), 0, ctx->stream, n, r64, b, norms);
    FSD_HIP(hipGetLastError());
    return 0;
}

int k_to64(fsd_ctx *ctx, int64_t n, const float *x, double *y)
{
    hipLaunchKernelGGL(to64_kernel, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, n, x, y);
    FSD_HIP(hipGetLastError());
    return 0;
}

int k_to32(fsd_ctx *ctx, int64_t n, const double *x, float *y)
{
    hipLaunchKernelGGL(to32_kernel, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, n, x, y);
    FSD_HIP(hipGetLastError());
    return 0;
}

int k_add64(fsd_ctx *ctx, int64_t n, const float *d, double *x)
{
    hipLaunchKernelGGL(add64_kernel, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, n, d, x);
    FSD_HIP(hipGetLastError());
    return 0;
}

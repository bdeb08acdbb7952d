// Saddle-point operator pieces, GMRES vector kernels (FP64 accumulation),
// right-hand-side assembly and the explicit Euler step with shear.
#include "fsd_internal.h"

#include <algorithm>

// ------------------------------------------------------------ vectors
__global__ void axpby_kernel(int64_t n, float a, const float *x, float b, const float *y, float *z)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = a * x[i] + (y ? b * y[i] : 0.f);
}

int k_axpby(fsd_ctx *ctx, int64_t n, float a, const float *x, float b, const float *y, float *z)
{
    hipLaunchKernelGGL(axpby_kernel, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, n, a, x, b, y, z);
    FSD_HIP(hipGetLastError());
    return 0;
}

__global__ void copy_kernel(int64_t n, const float *x, float *y)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i];
}

int k_copy(fsd_ctx *ctx, int64_t n, const float *x, float *y)
{
    hipLaunchKernelGGL(copy_kernel, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, n, x, y);
    FSD_HIP(hipGetLastError());
    return 0;
}

__global__ void fill_kernel(int64_t n, float v, float *y)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = v;
}

int k_fill(fsd_ctx *ctx, int64_t n, float v, float *y)
{
    hipLaunchKernelGGL(fill_kernel, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, n, v, y);
    FSD_HIP(hipGetLastError());
    return 0;
}

// partial dot products: block (chunk, k) reduces sum_q V_k[q] w[q] over its chunk
#define DOT_CHUNK 4096
__global__ void __launch_bounds__(256) multidot_partial_kernel(int64_t n, const float *V, int64_t ldv, const float *w,
                                                               double *part, int nchunks)
{
    __shared__ double red[4];
    const int k = blockIdx.y, ch = blockIdx.x;
    const float *v = V + (int64_t)k * ldv;
    const int64_t beg = (int64_t)ch * DOT_CHUNK;
    const int64_t end = min(n, beg + DOT_CHUNK);
    double s = 0.0;
    for (int64_t q = beg + threadIdx.x; q < end; q += 256) s += (double)v[q] * (double)w[q];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[(int64_t)k * nchunks + ch] = red[0] + red[1] + red[2] + red[3];
}

__global__ void __launch_bounds__(256) reduce_parts_kernel(const double *part, int nchunks, double *out)
{
    __shared__ double red[4];
    const int k = blockIdx.x;
    double s = 0.0;
    for (int c = threadIdx.x; c < nchunks; c += 256) s += part[(int64_t)k * nchunks + c];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[k] = red[0] + red[1] + red[2] + red[3];
}

// the same on float4 (n, ldv multiples of 4, 16-byte aligned vectors): a quarter of the load instructions
__global__ void __launch_bounds__(256) multidot_partial4_kernel(int64_t n4, const float4 *V, int64_t ldv4,
                                                                const float4 *w, double *part, int nchunks)
{
    __shared__ double red[4];
    const int k = blockIdx.y, ch = blockIdx.x;
    const float4 *v = V + (int64_t)k * ldv4;
    const int64_t beg = (int64_t)ch * (DOT_CHUNK / 4);
    const int64_t end = min(n4, beg + DOT_CHUNK / 4);
    double s = 0.0;
    for (int64_t q = beg + threadIdx.x; q < end; q += 256) {
        const float4 a = v[q], b = w[q];
        s += (double)a.x * (double)b.x + (double)a.y * (double)b.y + (double)a.z * (double)b.z +
             (double)a.w * (double)b.w;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[(int64_t)k * nchunks + ch] = red[0] + red[1] + red[2] + red[3];
}

static bool vec4_ok(int64_t n, int64_t ldv, const void *a, const void *b)
{
    return n % 4 == 0 && ldv % 4 == 0 && ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) == 0;
}

// out[k] = <V_k, w> for k < K (out is a device array of K doubles)
int k_multidot(fsd_ctx *ctx, int64_t n, int K, const float *V, int64_t ldv, const float *w, double *out)
{
    const int nchunks = (int)((n + DOT_CHUNK - 1) / DOT_CHUNK);
    if (vec4_ok(n, ldv, V, w))
        hipLaunchKernelGGL(multidot_partial4_kernel, dim3(nchunks, K), dim3(256), 0, ctx->stream, n / 4,
                           reinterpret_cast<const float4 *>(V), ldv / 4, reinterpret_cast<const float4 *>(w),
                           ctx->d_part, nchunks);
    else
        hipLaunchKernelGGL(multidot_partial_kernel, dim3(nchunks, K), dim3(256), 0, ctx->stream, n, V, ldv, w,
                           ctx->d_part, nchunks);
    hipLaunchKernelGGL(reduce_parts_kernel, dim3(K), dim3(256), 0, ctx->stream, ctx->d_part, nchunks, out);
    FSD_HIP(hipGetLastError());
    return 0;
}

// w -= sum_k h[k] V_k
__global__ void multi_axpy_kernel(int64_t n, int K, const float *V, int64_t ldv, const double *h, float *w)
{
    __shared__ float hs[256];
    for (int k = threadIdx.x; k < K; k += blockDim.x) hs[k] = (float)h[k];
    __syncthreads();
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    float s = w[q];
    for (int k = 0; k < K; ++k) s -= hs[k] * V[(int64_t)k * ldv + q];
    w[q] = s;
}

// the same on float4 (same order of operations per element)
__global__ void multi_axpy4_kernel(int64_t n4, int K, const float4 *V, int64_t ldv4, const double *h, float4 *w)
{
    __shared__ float hs[256];
    for (int k = threadIdx.x; k < K; k += blockDim.x) hs[k] = (float)h[k];
    __syncthreads();
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n4) return;
    float4 s = w[q];
    for (int k = 0; k < K; ++k) {
        const float4 v = V[(int64_t)k * ldv4 + q];
        s.x -= hs[k] * v.x; s.y -= hs[k] * v.y; s.z -= hs[k] * v.z; s.w -= hs[k] * v.w;
    }
    w[q] = s;
}

int k_multi_axpy(fsd_ctx *ctx, int64_t n, int K, const float *V, int64_t ldv, const double *h, float *w)
{
    if (K > 256) { ctx->err = "multi_axpy: more than 256 vectors"; return FSD_E_PARAM; }
    if (vec4_ok(n, ldv, V, w))
        hipLaunchKernelGGL(multi_axpy4_kernel, dim3(nblk(n / 4, 256)), dim3(256), 0, ctx->stream, n / 4, K,
                           reinterpret_cast<const float4 *>(V), ldv / 4, h, reinterpret_cast<float4 *>(w));
    else
        hipLaunchKernelGGL(multi_axpy_kernel, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, n, K, V, ldv, h, w);
    FSD_HIP(hipGetLastError());
    return 0;
}

// v *= 1 / hn, hn = sqrt(max(h[K] - sum_{k<K} h[k]^2, 0)): the norm of the second
// Gram-Schmidt pass's result from its dot products (h[0..K)) and |w'|^2 (h[K]).
// Every block forms hn the same way; block 0 stores it in *hn_out.
__global__ void __launch_bounds__(256) scale_by_norm_kernel(int64_t n, float *v, const double *h, int K,
                                                            double *hn_out)
{
    __shared__ float s_inv;
    if (threadIdx.x < 64) {
        double acc = 0.0;
        for (int k = threadIdx.x; k < K; k += 64) acc += h[k] * h[k];
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (threadIdx.x == 0) {
            const double hn = sqrt(fmax(h[K] - acc, 0.0));
            s_inv = hn > 0.0 ? (float)(1.0 / hn) : 0.f;
            if (blockIdx.x == 0) *hn_out = hn;
        }
    }
    __syncthreads();
    const float inv = s_inv;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        v[q] *= inv;
}

int k_scale_by_norm(fsd_ctx *ctx, int64_t n, float *v, const double *h, int K, double *hn_out)
{
    hipLaunchKernelGGL(scale_by_norm_kernel, dim3(std::min<int64_t>(nblk(n, 256), 1024)), dim3(256), 0, ctx->stream,
                       n, v, h, K, hn_out);
    FSD_HIP(hipGetLastError());
    return 0;
}

int k_nrm2(fsd_ctx *ctx, int64_t n, const float *x, double *out)
{
    return k_multidot(ctx, n, 1, x, 0, x, out);  // returns the squared norm
}

// ------------------------------------------------------------ saddle
// Saddle_Preconditioner (Saddle.cu:171-264):
//   x_F = b_F - G b_F + G b_U,  x_S = b_S,  x_U = G b_F - G b_U
__global__ void saddle_combine_kernel(int N, const float *b, const float *gF, const float *gU, float *x)
{
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n6 = 6 * (int64_t)N;
    if (q < n6) {
        const float f = gF[q], u = gU[q];
        x[q] = b[q] - f + u;
        x[11 * (int64_t)N + q] = f - u;
    } else if (q < 11 * (int64_t)N) {
        x[q] = b[q];
    }
}

int k_precond_apply2(fsd_ctx *ctx, const float *b0, const float *b1, float *y0, float *y1, int n_rhs);
int k_saddle_precond_tp(fsd_ctx *ctx, const float *b, float *x);

int k_saddle_precond(fsd_ctx *ctx, const float *b, float *x)
{
    if (ctx->precond_mode == 2 && lp_active(ctx)) return lp_apply(ctx, b, b + 11 * (int64_t)ctx->N, b, x, 1);
    if (ctx->ic_jac_k > 0) return k_saddle_precond_jacobi(ctx, b, x);
    // one sweep pair on d = b_F - b_U when the one-launch sweeps are built
    const int tp = k_saddle_precond_tp(ctx, b, x);
    if (tp <= 0) return tp;
    const int N = ctx->N;
    float *gF = ctx->d_g12, *gU = ctx->d_g12 + 6 * (int64_t)N;
    int rc = k_precond_apply2(ctx, b, b + 11 * (int64_t)N, gF, gU, 2);
    if (rc) return rc;
    hipLaunchKernelGGL(saddle_combine_kernel, dim3(nblk(11 * (int64_t)N, 256)), dim3(256), 0, ctx->stream, N, b, gF,
                       gU, x);
    FSD_HIP(hipGetLastError());
    return 0;
}

// Saddle_Multiply (Saddle.cu:80-151): y_top = M x_top + [x_U; 0], y_bot = x_F - R_FU x_U
int k_saddle_multiply(fsd_ctx *ctx, const float4 *pos, const float *x, float *y)
{
    const int64_t N = ctx->N;
    int rc = k_mobility(ctx, pos, x, y, 3, x + 11 * N);
    if (rc) return rc;
    return k_rfu_apply(ctx, x + 11 * N, y + 11 * N, x, -1.0f);
}

// ------------------------------------------------------------ RHS
// Integrator_AddStrainRate_kernel (Helper_Integrator.cu:114-158) and
// rhs[11N:] -= F_applied (Integrator.cu:831-841); rhs assumed zeroed.
__global__ void strain_rhs_kernel(int N, float *rhs, float shear_rate, float B2, const float *mask2, const float *ori,
                                  const float *F)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float *e = rhs + 6 * (int64_t)N + 5 * (int64_t)i;
    e[1] += shear_rate;
    if (B2 != 0.f && ori) {
        const float px = ori[3 * i], py = ori[3 * i + 1], pz = ori[3 * i + 2];
        const float b2 = -0.6f * B2 * mask2[i];
        e[0] -= b2 * (px * px - pz * pz);
        e[1] -= b2 * (2.f * px * py);
        e[2] -= b2 * (2.f * px * pz);
        e[3] -= b2 * (2.f * py * pz);
        e[4] -= b2 * (py * py - pz * pz);
    }
    float *u = rhs + 11 * (int64_t)N + 6 * (int64_t)i;
    for (int a = 0; a < 6; ++a) u[a] -= F[6 * (int64_t)i + a];
}

int k_strain_rhs(fsd_ctx *ctx, float *rhs, float shear_rate, const float *ori, const float *force)
{
    hipLaunchKernelGGL(strain_rhs_kernel, dim3(nblk(ctx->N, 256)), dim3(256), 0, ctx->stream, ctx->N, rhs, shear_rate,
                       ori ? ctx->p.sqm_B2 : 0.f, ctx->d_mask2, ori, force);
    FSD_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------ Euler
// Integrator_ExplicitEuler_Shear_kernel (Integrator.cu:126-228)
__global__ void euler_shear_kernel(int N, float4 *pos, float *ori, float4 *pos_gb, int *image, const float *vel,
                                   float B1, const float *mask1, const float *noise, BoxF b, float dt, float shear_rate)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float4 p4 = pos[i];
    float px = p4.x, py = p4.y, pz = p4.z;
    float vx = vel[6 * (int64_t)i], vy = vel[6 * (int64_t)i + 1], vz = vel[6 * (int64_t)i + 2];
    float wx = vel[6 * (int64_t)i + 3], wy = vel[6 * (int64_t)i + 4], wz = vel[6 * (int64_t)i + 5];
    vx += shear_rate * py;
    wz -= shear_rate / 2.f;
    if (noise) { wx += noise[3 * i]; wy += noise[3 * i + 1]; wz += noise[3 * i + 2]; }
    float q0 = 1.f, qx = 0.f, qy = 0.f, qz = 0.f;
    float aw = sqrtf(wx * wx + wy * wy + wz * wz);
    if (aw > 1e-6f) {
        float ang = aw * dt;
        q0 = cosf(ang / 2.f);
        float s = sinf(ang / 2.f) / aw;
        qx = wx * s; qy = wy * s; qz = wz * s;
    }
    float ox = 0.f, oy = 0.f, oz = 0.f;
    if (ori) { ox = ori[3 * i]; oy = ori[3 * i + 1]; oz = ori[3 * i + 2]; }
    const float m1 = mask1 ? mask1[i] : 1.f;
    const float act = -2.f / 3.f * B1 * m1;
    vx -= act * ox; vy -= act * oy; vz -= act * oz;
    float4 g = pos_gb[i];
    g.x += (vx - shear_rate * py + shear_rate * g.y) * dt;
    g.y += vy * dt;
    g.z += vz * dt;
    pos_gb[i] = g;
    px += vx * dt; py += vy * dt; pz += vz * dt;
    int ix = image[3 * i], iy = image[3 * i + 1], iz = image[3 * i + 2];
    box_wrap(b, px, py, pz, ix, iy, iz);
    pos[i] = make_float4(px, py, pz, p4.w);
    image[3 * i] = ix; image[3 * i + 1] = iy; image[3 * i + 2] = iz;
    if (ori) {
        float qq = qx * qx + qy * qy + qz * qz, qp = qx * ox + qy * oy + qz * oz;
        float cx = qy * oz - qz * oy, cy = qz * ox - qx * oz, cz = qx * oy - qy * ox;
        float nx = (q0 * q0 - qq) * ox + 2.f * qp * qx + 2.f * q0 * cx;
        float ny = (q0 * q0 - qq) * oy + 2.f * qp * qy + 2.f * q0 * cy;
        float nz = (q0 * q0 - qq) * oz + 2.f * qp * qz + 2.f * q0 * cz;
        float m = sqrtf(nx * nx + ny * ny + nz * nz);
        ori[3 * i] = nx / m; ori[3 * i + 1] = ny / m; ori[3 * i + 2] = nz / m;
    }
}

// Director rotation pv <- q pv q* with the quaternion of angular velocity w over
// time tau (Integrator.cu:278-286, 312-321), renormalised
__device__ inline void rotate_director(float wx, float wy, float wz, float tau, float &ox, float &oy, float &oz)
{
    float q0 = 1.f, qx = 0.f, qy = 0.f, qz = 0.f;
    const float aw = sqrtf(wx * wx + wy * wy + wz * wz);
    if (aw > 1e-6f) {
        const float ang = aw * tau;
        q0 = cosf(ang / 2.f);
        const float s = sinf(ang / 2.f) / aw;
        qx = wx * s; qy = wy * s; qz = wz * s;
    }
    const float qq = qx * qx + qy * qy + qz * qz, qp = qx * ox + qy * oy + qz * oz;
    const float cx = qy * oz - qz * oy, cy = qz * ox - qx * oz, cz = qx * oy - qy * ox;
    const float nx = (q0 * q0 - qq) * ox + 2.f * qp * qx + 2.f * q0 * cx;
    const float ny = (q0 * q0 - qq) * oy + 2.f * qp * qy + 2.f * q0 * cy;
    const float nz = (q0 * q0 - qq) * oz + 2.f * qp * qz + 2.f * q0 * cz;
    const float m = sqrtf(nx * nx + ny * ny + nz * nz);
    ox = nx / m; oy = ny / m; oz = nz / m;
}

// Integrator_ExplicitEuler1_Shear_kernel (Integrator.cu:232-326): first stage of
// the second-order Runge-Kutta scheme.  The unwrapped position advances half a
// step, the periodic position a full step without wrapping (the second stage
// averages it), pos.w is kept.
__global__ void rk2_stage1_kernel(int N, const float4 *pos_in, const float *ori_in, float4 *pos_out, float *ori_out,
                                  float4 *pos_gb, const float *vel, float B1, const float *mask1, const float *noise,
                                  float dt, float shear_rate)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float4 p4 = pos_in[i];
    float px = p4.x, py = p4.y, pz = p4.z;
    float vx = vel[6 * (int64_t)i], vy = vel[6 * (int64_t)i + 1], vz = vel[6 * (int64_t)i + 2];
    float wx = vel[6 * (int64_t)i + 3], wy = vel[6 * (int64_t)i + 4], wz = vel[6 * (int64_t)i + 5];
    vx += shear_rate * py;
    wz -= shear_rate / 2.f;
    if (noise) { wx += noise[3 * i]; wy += noise[3 * i + 1]; wz += noise[3 * i + 2]; }
    float ox = ori_in[3 * i], oy = ori_in[3 * i + 1], oz = ori_in[3 * i + 2];
    const float act = -2.f / 3.f * B1 * (mask1 ? mask1[i] : 1.f);
    vx -= act * ox; vy -= act * oy; vz -= act * oz;
    float4 g = pos_gb[i];
    g.x += (vx - shear_rate * py + shear_rate * g.y) * 0.5f * dt;
    g.y += vy * 0.5f * dt;
    g.z += vz * 0.5f * dt;
    pos_gb[i] = g;
    px += vx * dt; py += vy * dt; pz += vz * dt;
    pos_out[i] = make_float4(px, py, pz, p4.w);
    rotate_director(wx, wy, wz, dt, ox, oy, oz);
    ori_out[3 * i] = ox; ori_out[3 * i + 1] = oy; ori_out[3 * i + 2] = oz;
}

// Integrator_RK_Shear_kernel (Integrator.cu:329-438): second stage.  Position
// and director are the combinations c1 * state1 + c2 * state2 (the director
// renormalised), moved by c3 * dt and wrapped; the unwrapped position advances
// half a step; pos.w is written as 0 as the reference does.
__global__ void rk2_stage2_kernel(int N, float c1, const float4 *pos1, const float *ori1, float c2, const float4 *pos2,
                                  const float *ori2, float4 *pos_out, float *ori_out, float4 *pos_gb, const float *vel,
                                  float B1, const float *mask1, const float *noise, int *image, BoxF b, float c3,
                                  float dt, float shear_rate)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float4 a4 = pos1[i], b4 = pos2[i];
    float px = c1 * a4.x + c2 * b4.x, py = c1 * a4.y + c2 * b4.y, pz = c1 * a4.z + c2 * b4.z;
    float vx = vel[6 * (int64_t)i], vy = vel[6 * (int64_t)i + 1], vz = vel[6 * (int64_t)i + 2];
    float wx = vel[6 * (int64_t)i + 3], wy = vel[6 * (int64_t)i + 4], wz = vel[6 * (int64_t)i + 5];
    vx += shear_rate * py;
    wz -= shear_rate / 2.f;
    if (noise) { wx += noise[3 * i]; wy += noise[3 * i + 1]; wz += noise[3 * i + 2]; }
    float ox = ori1[3 * i] * c1 + ori2[3 * i] * c2;
    float oy = ori1[3 * i + 1] * c1 + ori2[3 * i + 1] * c2;
    float oz = ori1[3 * i + 2] * c1 + ori2[3 * i + 2] * c2;
    const float m0 = sqrtf(ox * ox + oy * oy + oz * oz);
    ox /= m0; oy /= m0; oz /= m0;
    const float act = -2.f / 3.f * B1 * (mask1 ? mask1[i] : 1.f);
    vx -= act * ox; vy -= act * oy; vz -= act * oz;
    float4 g = pos_gb[i];
    g.x += (vx - shear_rate * py + shear_rate * g.y) * 0.5f * dt;
    g.y += vy * 0.5f * dt;
    g.z += vz * 0.5f * dt;
    pos_gb[i] = g;
    px += vx * dt * c3; py += vy * dt * c3; pz += vz * dt * c3;
    int ix = image[3 * i], iy = image[3 * i + 1], iz = image[3 * i + 2];
    box_wrap(b, px, py, pz, ix, iy, iz);
    pos_out[i] = make_float4(px, py, pz, 0.f);
    image[3 * i] = ix; image[3 * i + 1] = iy; image[3 * i + 2] = iz;
    rotate_director(wx, wy, wz, dt * c3, ox, oy, oz);
    ori_out[3 * i] = ox; ori_out[3 * i + 1] = oy; ori_out[3 * i + 2] = oz;
}

int k_rk2_stage1(fsd_ctx *ctx, const float4 *pos_in, const float *ori_in, float4 *pos_out, float *ori_out,
                 float4 *pos_gb, const float *vel, const float *mask1, const float *noise, float dt, float shear_rate)
{
    hipLaunchKernelGGL(rk2_stage1_kernel, dim3(nblk(ctx->N, 256)), dim3(256), 0, ctx->stream, ctx->N, pos_in, ori_in,
                       pos_out, ori_out, pos_gb, vel, ctx->p.sqm_B1, mask1, noise, dt, shear_rate);
    FSD_HIP(hipGetLastError());
    return 0;
}

int k_rk2_stage2(fsd_ctx *ctx, float c1, const float4 *pos1, const float *ori1, float c2, const float4 *pos2,
                 const float *ori2, float4 *pos_out, float *ori_out, float4 *pos_gb, const float *vel,
                 const float *mask1, const float *noise, int *image, float c3, float dt, float shear_rate)
{
    hipLaunchKernelGGL(rk2_stage2_kernel, dim3(nblk(ctx->N, 256)), dim3(256), 0, ctx->stream, ctx->N, c1, pos1, ori1,
                       c2, pos2, ori2, pos_out, ori_out, pos_gb, vel, ctx->p.sqm_B1, mask1, noise, image, ctx->box, c3,
                       dt, shear_rate);
    FSD_HIP(hipGetLastError());
    return 0;
}

int k_euler_shear(fsd_ctx *ctx, float4 *pos, float *ori, float4 *pos_gb, int *image, const float *vel,
                  const float *mask1, const float *noise, float dt, float shear_rate, float B1)
{
    hipLaunchKernelGGL(euler_shear_kernel, dim3(nblk(ctx->N, 256)), dim3(256), 0, ctx->stream, ctx->N, pos, ori,
                       pos_gb, image, vel, B1, mask1, noise, ctx->box, dt, shear_rate);
    FSD_HIP(hipGetLastError());
    return 0;
}

// Brownian motion of the FSD step (T > 0): Gaussian noise, the wave-space
// part of the far-field Brownian slip, the random-finite-difference shift,
// and small vector helpers of the Lanczos square roots.
#include "fsd_internal.h"

// ----------------------------------------------------------- normal noise
// Counter-based Philox-4x32 with 10 rounds (Salmon, Moraes, Dror and Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11): counter
// (index, timestep, stream, 0), key (seed, 0).  Each counter gives four
// uniforms in (0, 1); Box-Muller turns each pair into two normals.
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(M0, c[0]), lo0 = M0 * c[0];
        const uint32_t hi1 = __umulhi(M1, c[2]), lo1 = M1 * c[2];
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += W0; k1 += W1;
    }
}

__global__ void normal_kernel(int64_t n, uint32_t seed, uint32_t step, uint32_t stream, float sd, float *out)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (4 * t >= n) return;
    uint32_t c[4] = {(uint32_t)t, step, stream, (uint32_t)(t >> 32)};
    philox4x32_10(c, seed, 0u);
    float z[4];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const float u1 = ((float)c[2 * p] + 0.5f) * 2.3283064365386963e-10f;      // (0, 1)
        const float u2 = ((float)c[2 * p + 1] + 0.5f) * 2.3283064365386963e-10f;
        const float r = sqrtf(-2.0f * logf(u1));
        float s, co;
        sincosf(6.283185307179586f * u2, &s, &co);
        z[2 * p] = r * co;
        z[2 * p + 1] = r * s;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (4 * t + q < n) out[4 * t + q] = sd * z[q];
}

int k_normals(fsd_ctx *ctx, int64_t n, uint32_t step, uint32_t stream, float sd, float *out)
{
    const int64_t nt = (n + 3) / 4;
    hipLaunchKernelGGL(normal_kernel, dim3(nblk(nt, 256)), dim3(256), 0, ctx->stream, n, ctx->p.seed, step, stream,
                       sd, out);
    FSD_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------- far field, wave-space part
// Brownian_FarField_RNG_Grid1of2 + Grid2of2 (Brownian_FarField.cu:113-395)
// written directly on the R2C half spectrum.  Grid1of2 writes, for every
// point of its "lower half" set S, the noise at the point and its conjugate
// at the conjugate point (real and scaled by sqrt 2 at the self-conjugate
// Nyquist corners); the value a half-spectrum point ends with is the one the
// highest-index writer in S gives it, which makes the grid exactly Hermitian
// (the reference writes from all threads at once and races where a point and
// its conjugate are both in S).  Noise of point idx is read at [q, q + 6),
// q = idx (parity: the reference's overlapping reads) or 6 idx.
__device__ __forceinline__ bool in_lower_half(int i, int j, int k, int Nx, int Ny, int Nz)
{
    if (2 * k >= Nz + 1) return false;
    if (k == 0 && 2 * j >= Ny + 1) return false;
    if (k == 0 && j == 0 && 2 * i >= Nx + 1) return false;
    return !(i == 0 && j == 0 && k == 0);
}

// Grid2of2: square root of the wave-space Green's function on one mode
__device__ inline void noise_scale(const float4 tk, bool origin, const float2 f[3], float2 o[11])
{
    const float ksq = tk.x * tk.x + tk.y * tk.y + tk.z * tk.z;
    const float kk = sqrtf(ksq);
    const float B = origin ? 0.f : sqrtf(tk.w);
    const float SU = origin ? 0.f : sinf(kk) / kk;
    const float SD = origin ? 0.f : 3.0f * (sinf(kk) - kk * cosf(kk)) / (ksq * kk);
    float2 kdF = make_float2(0.f, 0.f);
    if (!origin) {
        kdF.x = (tk.x * f[0].x + tk.y * f[1].x + tk.z * f[2].x) / ksq;
        kdF.y = (tk.x * f[0].y + tk.y * f[1].y + tk.z * f[2].y) / ksq;
    }
    const float kv[3] = {tk.x, tk.y, tk.z};
    float2 W[3];
    for (int a = 0; a < 3; ++a) {
        W[a].x = (f[a].x - kv[a] * kdF.x) * B;
        W[a].y = (f[a].y - kv[a] * kdF.y) * B;
    }
    for (int a = 0; a < 3; ++a) o[a] = make_float2(SU * W[a].x, SU * W[a].y);
    // gradient fields XX XY XZ YZ YY YX ZX ZY: component (a, c) = i SD W_a k_c
    const int sa[8] = {0, 0, 0, 1, 1, 1, 2, 2}, sc[8] = {0, 1, 2, 2, 1, 0, 0, 1};
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const float2 w = W[sa[s]];
        const float kc_ = kv[sc[s]];
        o[3 + s] = make_float2(-SD * w.y * kc_, SD * w.x * kc_);
    }
}

struct NoiseGeom {
    int Nx, Ny, Nz;
    BoxF b;
    float xi, eta, twopi_k;
};

__global__ void ff_noise_kernel(const float *noise, int parity, float fac, const float4 *gk, NoiseGeom g, int64_t Mh,
                                float2 *spec)
{
    const int64_t h = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (h >= Mh) return;
    const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
    const int Nzh = Nz / 2 + 1;
    const int k = (int)(h % Nzh), j = (int)((h / Nzh) % Ny), i = (int)(h / ((int64_t)Nzh * Ny));
    const bool in = (i == Nx / 2) && (Nx % 2 == 0), jn = (j == Ny / 2) && (Ny % 2 == 0),
               kn = (k == Nz / 2) && (Nz % 2 == 0);
    const int ic = (i == 0 || in) ? i : Nx - i, jc = (j == 0 || jn) ? j : Ny - j, kc = (k == 0 || kn) ? k : Nz - k;
    const int64_t idx = ((int64_t)i * Ny + j) * Nz + k, cidx = ((int64_t)ic * Ny + jc) * Nz + kc;
    const bool own = in_lower_half(i, j, k, Nx, Ny, Nz);
    const bool partner = (cidx != idx) && in_lower_half(ic, jc, kc, Nx, Ny, Nz);
    float2 f[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
    if (own && (!partner || idx > cidx)) {
        const int64_t q = parity ? idx : 6 * idx;
        const bool special = (i == 0 || in) && (j == 0 || jn) && (k == 0 || kn);
        for (int a = 0; a < 3; ++a)
            f[a] = special ? make_float2(1.414213562373095f * fac * noise[q + a], 0.f)
                           : make_float2(fac * noise[q + a], fac * noise[q + 3 + a]);
    } else if (partner) {
        const int64_t q = parity ? cidx : 6 * cidx;
        for (int a = 0; a < 3; ++a) f[a] = make_float2(fac * noise[q + a], -fac * noise[q + 3 + a]);
    }
    float2 o[11];
    noise_scale(gk[h], h == 0, f, o);
    if (has_nyquist(i, j, k, Nx, Ny, Nz)) {
        // the conjugate mode holds conj(f) scaled with its own wave vector
        const float4 tkc = gridk_value(ic, jc, kc, Nx, Ny, Nz, g.b, g.xi, g.eta, g.twopi_k);
        const float2 fc[3] = {make_float2(f[0].x, -f[0].y), make_float2(f[1].x, -f[1].y),
                              make_float2(f[2].x, -f[2].y)};
        float2 oc[11];
        noise_scale(tkc, false, fc, oc);
#pragma unroll
        for (int c = 0; c < 11; ++c) o[c] = make_float2(0.5f * (o[c].x + oc[c].x), 0.5f * (o[c].y - oc[c].y));
    }
#pragma unroll
    for (int c = 0; c < 11; ++c) spec[c * Mh + h] = o[c];
}

int k_ff_noise(fsd_ctx *ctx, const float *noise, float fac)
{
    const fsd_pse_info &ps = ctx->pse;
    NoiseGeom g{ps.Nx, ps.Ny, ps.Nz, ctx->box, ctx->p.xi, ps.eta, k_twopi(ctx)};
    hipLaunchKernelGGL(ff_noise_kernel, dim3(nblk(ctx->Mh, 256)), dim3(256), 0, ctx->stream, noise, ctx->p.parity,
                       fac, ctx->d_gridk, g, (int64_t)ctx->Mh, ctx->d_spec);
    FSD_HIP(hipGetLastError());
    return 0;
}

// wave-space contraction result (u, gradient in slot order) -> generalised
// 11N velocity [U, W | E5], added to out (Mobility_D2WE_kernel,
// Saddle_MakeGeneralizedU_kernel)
__global__ void wave_to_general_kernel(int N, const float *wave, float *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float *w = wave + 11 * (int64_t)i;
    const float *D = w + 3;
    const float W0 = 0.5f * (D[3] - D[7]), W1 = 0.5f * (D[6] - D[2]), W2 = 0.5f * (D[1] - D[5]);
    const float e0 = D[0], e1 = 0.5f * (D[1] + D[5]), e2 = 0.5f * (D[2] + D[6]), e3 = 0.5f * (D[3] + D[7]), e4 = D[4];
    float *o6 = out + 6 * (int64_t)i;
    o6[0] += w[0]; o6[1] += w[1]; o6[2] += w[2];
    o6[3] += W0; o6[4] += W1; o6[5] += W2;
    float *o5 = out + 6 * (int64_t)N + 5 * (int64_t)i;
    o5[0] += 2.f * e0 + e4; o5[1] += 2.f * e1; o5[2] += 2.f * e2; o5[3] += 2.f * e3; o5[4] += 2.f * e4 + e0;
}

int k_wave_to_general(fsd_ctx *ctx, float *out)
{
    hipLaunchKernelGGL(wave_to_general_kernel, dim3(nblk(ctx->N, 256)), dim3(256), 0, ctx->stream, ctx->N,
                       ctx->d_wave, out);
    FSD_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------------- RFD
// Integrator_ExplicitEuler_kernel (Integrator.cu:61-93) as Integrator_RFD uses
// it: pos' = wrap(pos + scale psi_xyz), the image kept local.
__global__ void shift_wrap_kernel(int N, const float4 *pos, const float *psi6, BoxF b, float scale, float4 *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float4 p = pos[i];
    float x = p.x + scale * psi6[6 * (int64_t)i], y = p.y + scale * psi6[6 * (int64_t)i + 1],
          z = p.z + scale * psi6[6 * (int64_t)i + 2];
    int ix = 0, iy = 0, iz = 0;
    box_wrap(b, x, y, z, ix, iy, iz);
    out[i] = make_float4(x, y, z, p.w);
}

int k_shift_wrap(fsd_ctx *ctx, const float4 *pos, const float *psi6, float scale, float4 *out)
{
    hipLaunchKernelGGL(shift_wrap_kernel, dim3(nblk(ctx->N, 256)), dim3(256), 0, ctx->stream, ctx->N, pos, psi6,
                       ctx->box, scale, out);
    FSD_HIP(hipGetLastError());
    return 0;
}

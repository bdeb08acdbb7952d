// Learned preconditioner factor (learned.h): features, the BF16-MFMA
// message-passing network, factor assembly with the pivot check, the factor's
// application by block Jacobi iterations, and the training loss with its
// backward pass (own kernels throughout; no library GEMM).
//
// It replaces the IC(0) factorisation of the saddle preconditioner
// (Precondition.cu:1116-1307, 1448-1489) on the pattern that
// Precondition.cu:530-836 builds; the pattern, the RCM permutation and the
// pair blocks A are this library's IC(0) path's own (k_precond.hip).
//
// Per-row vectors of the factor are [N][6] in permuted order.  Row kernels
// give six lanes to a row (component a) and ten rows to a 64-lane wave.
#include "fsd_internal.h"
#include "learned.h"

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------ layer table
void lp::layers(Layer out[NLAYER], int64_t *nparam, int64_t *nbf)
{
    const int dims[NLAYER][3] = {
        {FIN, H, 1}, {FIN, H, 1},                        // particle and edge encoders
        {3 * H, H, 1}, {2 * H, H, 1},                    // layer 1: message, particle update
        {3 * H, H, 1}, {2 * H, H, 1},                    // layer 2
        {3 * H, H, 1}, {2 * H, H, 1},                    // layer 3
        {3 * H, OE, 0}, {H, ON, 0}};                     // edge and particle heads (linear)
    int64_t off = 0, boff = 0;
    for (int l = 0; l < NLAYER; ++l) {
        out[l].ki = dims[l][0];
        out[l].no = dims[l][1];
        out[l].act = dims[l][2];
        out[l].w_off = off;
        off += (int64_t)out[l].ki * out[l].no;
        out[l].b_off = off;
        off += out[l].no;
        out[l].bf_off = boff;
        boff += 2 * (int64_t)out[l].ki * out[l].no;
    }
    *nparam = off;
    *nbf = boff;
}

namespace {
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float silu_d(float z)
{
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ float sgnlog(float v) { return copysignf(log1pf(fabsf(v)), v); }
// the network's raw outputs enter the factor bounded: pair-coefficient multipliers
// 1 + tanh(c) in (0, 2), fill weights 2 sigmoid(c + c0) in (0, 2) (0.98 at c = 0: the
// zero-output factor subtracts nearly the whole one-step fill term; above 1 the fill
// can reach IC(0)'s, whose pivots D_Q are below A_QQ + lambda S), pivot shift
// lambda e^a S (a < 0 shrinks the pivots towards IC(0)'s)
constexpr float LP_GAMMA0 = -0.036004f;   // logit(0.491)
__device__ __forceinline__ float lp_gamma(float c) { return 2.0f * sigm(c + LP_GAMMA0); }
__device__ __forceinline__ float lp_gamma_d(float c)
{
    const float sg = sigm(c + LP_GAMMA0);
    return 2.0f * sg * (1.0f - sg);
}

__global__ void lp_fill_kernel(int64_t n, float v, float *p)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void lp_fill_int_kernel(int n, int v, int *p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// 6x6 symmetric positive definite inverse through Cholesky (lower triangle of D
// read); false on a non-positive pivot
__device__ bool chol6_inv(const float D[36], float Dinv[36])
{
    float Lm[36];
    for (int q = 0; q < 36; ++q) Lm[q] = 0.f;
    for (int c = 0; c < 6; ++c) {
        float s = D[c * 6 + c];
        for (int q = 0; q < c; ++q) s -= Lm[c * 6 + q] * Lm[c * 6 + q];
        if (!(s > 0.f)) return false;
        const float d = sqrtf(s);
        Lm[c * 6 + c] = d;
        for (int r = c + 1; r < 6; ++r) {
            float v = D[r * 6 + c];
            for (int q = 0; q < c; ++q) v -= Lm[r * 6 + q] * Lm[c * 6 + q];
            Lm[r * 6 + c] = v / d;
        }
    }
    // Y = L^-1 (lower), then Dinv = Y^T Y
    float Y[36];
    for (int q = 0; q < 36; ++q) Y[q] = 0.f;
    for (int c = 0; c < 6; ++c) {
        Y[c * 6 + c] = 1.0f / Lm[c * 6 + c];
        for (int r = c + 1; r < 6; ++r) {
            float s = 0.f;
            for (int q = c; q < r; ++q) s += Lm[r * 6 + q] * Y[q * 6 + c];
            Y[r * 6 + c] = -s / Lm[r * 6 + r];
        }
    }
    for (int a = 0; a < 6; ++a)
        for (int b = a; b < 6; ++b) {
            float s = 0.f;
            for (int q = b; q < 6; ++q) s += Y[q * 6 + a] * Y[q * 6 + b];
            Dinv[a * 6 + b] = s;
            Dinv[b * 6 + a] = s;
        }
    return true;
}

// eps.r as the pair blocks use it (k_precond.hip blocks_kernel)
__device__ __forceinline__ void eps_r(const float r[3], float E[3][3])
{
    E[0][0] = 0.f; E[0][1] = r[2]; E[0][2] = -r[1];
    E[1][0] = -r[2]; E[1][1] = 0.f; E[1][2] = r[0];
    E[2][0] = r[1]; E[2][1] = -r[0]; E[2][2] = 0.f;
}

// basis block k (0..5) of a pair with unit vector r and coefficients
// (XA12, YA12, YB12, XC12, YC12), entry (a, c) of the 6x6 block
__device__ __forceinline__ float basis_entry(int k, int a, int c, const float r[3], const float E[3][3],
                                             const float ph[5])
{
    const int ba = a / 3, bc = c / 3, ia = a % 3, ic = c % 3;
    const float rr = r[ia] * r[ic], del = (ia == ic) ? 1.f : 0.f;
    switch (k) {
    case 0: return (ba == 0 && bc == 0) ? ph[0] * rr : 0.f;
    case 1: return (ba == 0 && bc == 0) ? ph[1] * (del - rr) : 0.f;
    case 2: return (ba == 0 && bc == 1) ? ph[2] * E[ia][ic] : 0.f;
    case 3: return (ba == 1 && bc == 0) ? ph[2] * E[ia][ic] : 0.f;
    case 4: return (ba == 1 && bc == 1) ? ph[3] * rr : 0.f;
    default: return (ba == 1 && bc == 1) ? ph[4] * (del - rr) : 0.f;
    }
}

// six lanes per row, ten rows per wave
__device__ __forceinline__ bool row_lane(int N, int &P, int &a)
{
    // rows in one contiguous range per XCD: the blocks and iterate entries that neighbouring rows
    // share a cache line with are then read through the same L2
    const int lane = threadIdx.x & 63;
    const int wave = (int)((xcd_remap(blockIdx.x, gridDim.x) * (int64_t)blockDim.x + threadIdx.x) >> 6);
    const int g = lane / 6;
    a = lane - 6 * g;
    P = wave * 10 + g;
    return g < 10 && P < N;
}
inline int row_blocks(int N) { return nblk(nblk(N, 10) * 64, 256); }
// the six components of a row's vector from the row's six lanes (every lane of the wave must call)
__device__ __forceinline__ void row_gather(float s, float v[6])
{
    const int base = ((threadIdx.x & 63) / 6) * 6;
#pragma unroll
    for (int c = 0; c < 6; ++c) v[c] = __shfl(s, base + c);
}
}  // namespace

// ----------------------------------------------------------------- features
// row particle of every lower block
__global__ void lp_erow_kernel(int N, const int *lrow, int *erow)
{
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= N) return;
    for (int e = lrow[P]; e < lrow[P + 1]; ++e) erow[e] = P;
}

// per lower block: unit vector and pair coefficients (the values the pair block A
// uses, contact clamp at table row 0), and the edge input features
__global__ void lp_edge_feat_kernel(int64_t E, const int *erow, const int *lcol, const int *perm, const float4 *pos,
                                    BoxF b, LubTable t, float *geo, uint16_t *xe0)
{
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int i = perm[erow[e]], j = perm[lcol[e]];
    const float4 pi = pos[i], pj = pos[j];
    float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
    box_min_image(b, dx, dy, dz);
    const float dist = sqrtf(dist2_nc(dx, dy, dz));
    float v[10];
    if ((double)dist <= 2.0 + (double)t.table_min) {
        for (int c = 0; c < 10; ++c) v[c] = t.vals[c];
    } else {
        lub_lookup(t, dist, 0, 10, -1, v);
    }
    float *g = geo + e * 8;
    g[0] = dx / dist; g[1] = dy / dist; g[2] = dz / dist;
    g[3] = v[1]; g[4] = v[3]; g[5] = v[5]; g[6] = v[7]; g[7] = v[9];
    const float h = fmaxf(dist - 2.0f, 1.0e-4f);
    uint16_t *x = xe0 + e * lp::FIN;
    x[0] = f2bf(0.5f * log10f(h));
    for (int c = 0; c < 10; ++c) x[1 + c] = f2bf(0.25f * sgnlog(v[c]));
    for (int c = 11; c < lp::FIN; ++c) x[c] = 0;
}

// per particle: input features and the inverse of the uncorrected pivot A_PP + lambda S
__global__ void lp_node_feat_kernel(int N, const int *lrow, const int *urow, const float *Adiag, float lambda,
                                    uint16_t *xn0, float *Dtinv, int *fail)
{
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= N) return;
    const float *A = Adiag + (int64_t)P * 36;
    const float trU = A[0] + A[7] + A[14], trW = A[21] + A[28] + A[35];
    uint16_t *x = xn0 + (int64_t)P * lp::FIN;
    x[0] = f2bf(0.25f * (float)(lrow[P + 1] - lrow[P]));
    x[1] = f2bf(0.25f * (float)(urow[P + 1] - urow[P]));
    x[2] = f2bf(0.25f * log1pf(fmaxf(trU, 0.f) / 3.0f));
    x[3] = f2bf(0.25f * log1pf(fmaxf(trW, 0.f) / 3.0f));
    for (int c = 4; c < lp::FIN; ++c) x[c] = 0;
    float D[36];
    for (int q = 0; q < 36; ++q) D[q] = A[q];
    for (int c = 0; c < 6; ++c) D[7 * c] += lambda * (c < 3 ? 1.0f : 4.0f / 3.0f);
    float Di[36];
    if (!chol6_inv(D, Di)) {
        atomicOr(fail, 2);
        for (int q = 0; q < 36; ++q) Di[q] = 0.f;
    }
    for (int q = 0; q < 36; ++q) Dtinv[(int64_t)P * 36 + q] = Di[q];
}

// ------------------------------------------------------------ MFMA products
// Y[R x NO] = X[R x KI] B + bias with B given transposed (Bt [NO][KI]), BF16
// operands, FP32 accumulation (v_mfma_f32_16x16x32_bf16).  One wave per 16 rows:
// lane l feeds row (l & 15) of X and column (l & 15) of B with 8 consecutive k;
// its accumulator holds rows 4 (l >> 4) .. + 3 of that column.  Outputs
// (each optional): Z = pre-activation, A = act(Z), Y = res + act(Z).
template <int KI, int NO, int ACT>
__global__ void __launch_bounds__(256) lp_mm_kernel(int R, const uint16_t *__restrict__ X,
                                                    const uint16_t *__restrict__ Bt, const float *__restrict__ bias,
                                                    const float *__restrict__ res, float *__restrict__ Z,
                                                    float *__restrict__ A, float *__restrict__ Y)
{
    static_assert(KI % 32 == 0 && NO % 16 == 0, "tile sizes");
    const int lane = threadIdx.x & 63;
    const int r0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (r0 >= R) return;   // uniform across the wave
    const int arow = r0 + (lane & 15);
    const int kq = (lane >> 4) * 8;
    f32x4_t acc[NO / 16];
#pragma unroll
    for (int c = 0; c < NO / 16; ++c) acc[c] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < KI; kb += 32) {
        u32x4_t av = u32x4_t{0u, 0u, 0u, 0u};
        if (arow < R) av = *reinterpret_cast<const u32x4_t *>(X + (int64_t)arow * KI + kb + kq);
        const bf16x8_t a = __builtin_bit_cast(bf16x8_t, av);
#pragma unroll
        for (int cb = 0; cb < NO / 16; ++cb) {
            const u32x4_t bv = *reinterpret_cast<const u32x4_t *>(Bt + (int64_t)(cb * 16 + (lane & 15)) * KI + kb + kq);
            acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(bf16x8_t, bv), acc[cb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int cb = 0; cb < NO / 16; ++cb) {
        const int col = cb * 16 + (lane & 15);
        const float bb = bias ? bias[col] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = r0 + (lane >> 4) * 4 + i;
            if (row >= R) continue;
            const int64_t o = (int64_t)row * NO + col;
            const float z = acc[cb][i] + bb;
            const float y = ACT ? silu(z) : z;
            if (Z) Z[o] = z;
            if (A) A[o] = y;
            if (Y) Y[o] = (res ? res[o] : 0.f) + y;
        }
    }
}

// weight gradient partials: part[chunk] = X_chunk^T G_chunk (KI x NO) and the
// column sums of G_chunk (bias), over chunks of 256 rows staged 32 at a time
template <int KI, int NO>
__global__ void __launch_bounds__(256) lp_dw_kernel(int R, const uint16_t *__restrict__ X, const float *__restrict__ G,
                                                    float *__restrict__ part)
{
    constexpr int KG = 256 / NO, NQ = KI / KG, SUB = 32;
    __shared__ uint16_t Xs[SUB * KI];
    __shared__ float Gs[SUB * NO];
    const int t = threadIdx.x, n = t % NO, kg = t / NO;
    float acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.f;
    float accb = 0.f;
    const int rbeg = blockIdx.x * 256;
    for (int s = 0; s < 256; s += SUB) {
        const int rb = rbeg + s;
        for (int q = t; q < SUB * KI; q += 256) {
            const int r = rb + q / KI;
            Xs[q] = (r < R) ? X[(int64_t)r * KI + q % KI] : (uint16_t)0;
        }
        for (int q = t; q < SUB * NO; q += 256) {
            const int r = rb + q / NO;
            Gs[q] = (r < R) ? G[(int64_t)r * NO + q % NO] : 0.f;
        }
        __syncthreads();
        for (int r = 0; r < SUB; ++r) {
            const float g = Gs[r * NO + n];
#pragma unroll
            for (int q = 0; q < NQ; ++q) acc[q] += bf2f(Xs[r * KI + kg + KG * q]) * g;
            accb += g;
        }
        __syncthreads();
    }
    float *o = part + (int64_t)blockIdx.x * (KI * NO + NO);
#pragma unroll
    for (int q = 0; q < NQ; ++q) o[(kg + KG * q) * NO + n] = acc[q];
    if (kg == 0) o[KI * NO + n] = accb;
}

__global__ void lp_dw_reduce_kernel(int nchunk, int kn, int no, const float *part, float *gw, float *gb)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= kn + no) return;
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += part[(int64_t)c * (kn + no) + q];
    if (q < kn) gw[q] = s;
    else gb[q - kn] = s;
}

// ------------------------------------------------------ message passing glue
// [bf16 h_P | bf16 h_Q | bf16 e] per edge (P the row, Q the column)
__global__ void lp_gather_msg_kernel(int64_t E, const int *erow, const int *lcol, const float *h, const float *es,
                                     uint16_t *xm)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= E * 3 * lp::H) return;
    const int64_t e = t / (3 * lp::H);
    const int c = (int)(t - e * 3 * lp::H);
    float v;
    if (c < lp::H) v = h[(int64_t)erow[e] * lp::H + c];
    else if (c < 2 * lp::H) v = h[(int64_t)lcol[e] * lp::H + c - lp::H];
    else v = es[e * lp::H + c - 2 * lp::H];
    xm[t] = f2bf(v);
}

// messages summed at both end particles; writes [bf16 h | bf16 agg]
__global__ void lp_agg_kernel(int N, const int *lrow, const int *urow, const int *upos, const float *m, const float *h,
                              float *agg, uint16_t *xu)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * lp::H) return;
    const int P = (int)(t / lp::H), c = (int)(t % lp::H);
    float s = 0.f;
    for (int e = lrow[P]; e < lrow[P + 1]; ++e) s += m[(int64_t)e * lp::H + c];
    for (int u = urow[P]; u < urow[P + 1]; ++u) s += m[(int64_t)upos[u] * lp::H + c];
    agg[t] = s;
    xu[(int64_t)P * 2 * lp::H + c] = f2bf(h[t]);
    xu[(int64_t)P * 2 * lp::H + lp::H + c] = f2bf(s);
}

__global__ void lp_to_bf_kernel(int64_t n, const float *x, uint16_t *y)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// ------------------------------------------------------------ factor assembly
// W_e = sum_k (1 + tanh c_k) basis_k (the pair block A when every c_k = 0)
__global__ void lp_edge_W_kernel(int64_t E, const float *geo, const float *oute, float *W)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= E * 36) return;
    const int64_t e = t / 36;
    const int q = (int)(t - e * 36), a = q / 6, c = q % 6;
    const float *g = geo + e * 8;
    const float r[3] = {g[0], g[1], g[2]};
    const float ph[5] = {g[3], g[4], g[5], g[6], g[7]};
    float E3[3][3];
    eps_r(r, E3);
    const float *o = oute + e * lp::OE;
    float w = 0.f;
#pragma unroll
    for (int k = 0; k < 6; ++k) w += (1.0f + tanhf(o[k])) * basis_entry(k, a, c, r, E3, ph);
    W[t] = w;
}

// D_P = A_PP + lambda diag(e^a_t, e^a_r) S - s_P sum_{Q < P} gamma_e A_e F_Q A_e^T with
// gamma_e = 2 sigmoid(c_e + c0), A_e the pair block and F_Q the inverse pivot the fill uses:
// Dt_Q^-1 = (A_QQ + lambda S)^-1 in the first pass, the previous pass's D_Q^-1 in each later
// one (parallel sweeps of IC(0)'s sequential pivot recurrence; lp::NPASS of them).  The pivot is
// positive definite by construction: s_P is the first of 1, 1/2, 1/4, 1/8 for which the
// Cholesky factorisation of D_P succeeds (level 0 .. 3); when none does, D_P = Dt_P
// (level 4, with Dtinv its inverse), positive definite by the check of lp_node_feat_kernel.  lvl[P] records the
// level for the backward pass; fail[1] counts particles above level 0, fail[2] those at
// level 4.  Non-finite network outputs set fail bit 1 (the factor is then not used).
// The fill term uses the pair block, not W: the pivots then do not depend on W, and
// (D + W) D^-1 (D + W)^T is positive definite for any W once D is.
constexpr int LP_LEVELS = 4;
__global__ void lp_node_D_kernel(int N, const int *lrow, const int *lcol, const float *Adiag, float lambda,
                                 const float *outn, const float *oute, const float *Ablk, const float *Dtinv,
                                 const float *Finv, float *Dinv, float *Dout, int *lvl, int *fail)
{
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= N) return;
    float B[36], F[36];
    for (int q = 0; q < 36; ++q) {
        B[q] = Adiag[(int64_t)P * 36 + q];
        F[q] = 0.f;
    }
    const float at = expf(outn[(int64_t)P * lp::ON]), ar = expf(outn[(int64_t)P * lp::ON + 1]);
    bool finite = isfinite(at) && isfinite(ar);
    for (int c = 0; c < 6; ++c) B[7 * c] += lambda * (c < 3 ? at : ar * (4.0f / 3.0f));
    for (int e = lrow[P]; e < lrow[P + 1]; ++e) {
        const float *o = oute + (int64_t)e * lp::OE;
        const float gam = lp_gamma(o[6]);
        for (int k = 0; k < 7; ++k) finite = finite && isfinite(o[k]);
        const float *We = Ablk + (int64_t)e * 36, *Ti = Finv + (int64_t)lcol[e] * 36;
        float T[36];   // A F
        for (int a = 0; a < 6; ++a)
            for (int c = 0; c < 6; ++c) {
                float s = 0.f;
                for (int q = 0; q < 6; ++q) s += We[a * 6 + q] * Ti[q * 6 + c];
                T[a * 6 + c] = s;
            }
        for (int a = 0; a < 6; ++a)
            for (int c = 0; c < 6; ++c) {
                float s = 0.f;
                for (int q = 0; q < 6; ++q) s += T[a * 6 + q] * We[c * 6 + q];
                F[a * 6 + c] += gam * s;
            }
    }
    float D[36], Di[36];
    int level = LP_LEVELS;
    float sc = 1.0f;
    if (finite)
        for (int j = 0; j < LP_LEVELS; ++j, sc *= 0.5f) {
            for (int q = 0; q < 36; ++q) D[q] = B[q] - sc * F[q];
            if (chol6_inv(D, Di)) {
                level = j;
                break;
            }
        }
    else
        atomicOr(fail, 1);
    if (level == LP_LEVELS) {   // the uncorrected pivot A_PP + lambda S
        for (int q = 0; q < 36; ++q) D[q] = Adiag[(int64_t)P * 36 + q];
        for (int c = 0; c < 6; ++c) D[7 * c] += lambda * (c < 3 ? 1.0f : 4.0f / 3.0f);
        for (int q = 0; q < 36; ++q) Di[q] = Dtinv[(int64_t)P * 36 + q];
        atomicAdd(fail + 2, 1);
    }
    if (level > 0) atomicAdd(fail + 1, 1);
    lvl[P] = level;
    for (int q = 0; q < 36; ++q) Dinv[(int64_t)P * 36 + q] = Di[q];
    if (Dout)   // the pivot as used (lower triangle mirrored)
        for (int a = 0; a < 6; ++a)
            for (int c = 0; c < 6; ++c) Dout[(int64_t)P * 36 + a * 6 + c] = (c <= a) ? D[a * 6 + c] : D[c * 6 + a];
}

// --------------------------------------------------- block Jacobi application
// v_0 = Dinv t with t = (b0 - b1)[perm P]; t is stored when tout != null
__global__ void lj_pre_kernel(int N, const int *perm, const float *Dinv, const float *b0, const float *b1, float *tout,
                              float *vout)
{
    int P, a;
    const bool ok = row_lane(N, P, a);
    float tv = 0.f;
    if (ok) {
        const int64_t i = perm[P];
        tv = b0[6 * i + a];
        if (b1) tv -= b1[6 * i + a];
    }
    float t6[6];
    row_gather(tv, t6);
    if (!ok) return;
    const float *Dp = Dinv + (int64_t)P * 36 + 6 * a;
    float v = 0.f;
#pragma unroll
    for (int c = 0; c < 6; ++c) v += Dp[c] * t6[c];
    if (tout) tout[6 * (int64_t)P + a] = tv;
    vout[6 * (int64_t)P + a] = v;
}

// forward iteration: u = t - W v_in (stored when uout != null), v_out = Dinv u
__global__ void __launch_bounds__(256) lj_fwd_kernel(int N, const int *lrow, const int *lcol, const float *W,
                                                     const float *Dinv, const float *t, const float *vin, float *uout,
                                                     float *vout)
{
    int P, a;
    const bool ok = row_lane(N, P, a);
    float s = 0.f;
    if (ok) {
        s = t[6 * (int64_t)P + a];
        const int e1 = lrow[P + 1];
        for (int e = lrow[P]; e < e1; ++e) {
            const float *w = W + (int64_t)e * 36 + 6 * a;
            const float *v = vin + 6 * (int64_t)lcol[e];
            s -= w[0] * v[0] + w[1] * v[1] + w[2] * v[2] + w[3] * v[3] + w[4] * v[4] + w[5] * v[5];
        }
    }
    float s6[6];
    row_gather(s, s6);
    if (!ok) return;
    const float *Dp = Dinv + (int64_t)P * 36 + 6 * a;
    float v = 0.f;
#pragma unroll
    for (int c = 0; c < 6; ++c) v += Dp[c] * s6[c];
    if (uout) uout[6 * (int64_t)P + a] = s;
    vout[6 * (int64_t)P + a] = v;
}

// the preconditioner's output of a row's value g (unpermuted): mode 0: out = g (6N);
// mode 1: the saddle preconditioner's combination x_F = b_F - g, x_S = b_S, x_U = g
// (Saddle.cu:171-264 with g = G (b_F - b_U))
__device__ __forceinline__ void lj_post_write(int N, int P, int a, const int *perm, float g, const float *b, float *out,
                                              int mode)
{
    const int64_t i = perm[P], n6 = 6 * (int64_t)N, n11 = 11 * (int64_t)N;
    if (mode == 0) {
        out[6 * i + a] = g;
        return;
    }
    out[6 * i + a] = b[6 * i + a] - g;
    out[n11 + 6 * i + a] = g;
    if (a < 5) out[n6 + 5 * i + a] = b[n6 + 5 * i + a];
}

// backward iteration: s = W^T x_in (stored when sout != null), x_out = v_k - Dinv s; the last
// iteration of an application also writes the preconditioner's output (post_mode 0 or 1,
// lj_post_write; -1: none)
__global__ void __launch_bounds__(256) lj_bwd_kernel(int N, const int *urow, const int *ucol, const int *upos,
                                                     const float *W, const float *Dinv, const float *vk,
                                                     const float *xin, float *sout, float *xout,
                                                     const int *perm = nullptr, const float *bpost = nullptr,
                                                     float *outpost = nullptr, int post_mode = -1)
{
    int P, a;
    const bool ok = row_lane(N, P, a);
    float s = 0.f;
    if (ok) {
        const int e1 = urow[P + 1];
        for (int e = urow[P]; e < e1; ++e) {
            const float *w = W + (int64_t)upos[e] * 36 + a;
            const float *x = xin + 6 * (int64_t)ucol[e];
            s += w[0] * x[0] + w[6] * x[1] + w[12] * x[2] + w[18] * x[3] + w[24] * x[4] + w[30] * x[5];
        }
    }
    float s6[6];
    row_gather(s, s6);
    if (!ok) return;
    const float *Dp = Dinv + (int64_t)P * 36 + 6 * a;
    float q = 0.f;
#pragma unroll
    for (int c = 0; c < 6; ++c) q += Dp[c] * s6[c];
    if (sout) sout[6 * (int64_t)P + a] = s;
    const float x = vk[6 * (int64_t)P + a] - q;
    xout[6 * (int64_t)P + a] = x;
    if (post_mode >= 0) lj_post_write(N, P, a, perm, x, bpost, outpost, post_mode);
}

// g = x unpermuted; mode 0: out = g (6N); mode 1: the saddle preconditioner's
// combination x_F = b_F - g, x_S = b_S, x_U = g (Saddle.cu:171-264 with g = G (b_F - b_U))
__global__ void lj_post_kernel(int N, const int *perm, const float *x, const float *b, float *out, int mode)
{
    int P, a;
    if (!row_lane(N, P, a)) return;
    lj_post_write(N, P, a, perm, x[6 * (int64_t)P + a], b, out, mode);
}

// ---------------------------------------------------- training loss, adjoint
// out_P = scale Dinv_P in_P
__global__ void lj_dmul_kernel(int N, const float *Dinv, const float *in, float scale, float *out)
{
    int P, a;
    const bool ok = row_lane(N, P, a);
    const float iv = ok ? in[6 * (int64_t)P + a] : 0.f;
    float v6[6];
    row_gather(iv, v6);
    if (!ok) return;
    const float *Dp = Dinv + (int64_t)P * 36 + 6 * a;
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < 6; ++c) s += Dp[c] * v6[c];
    out[6 * (int64_t)P + a] = scale * s;
}

// out_P = sign sum_{e = (P, Q)} W_e in_Q   (lower product)
__global__ void lj_lower_kernel(int N, const int *lrow, const int *lcol, const float *W, const float *in, float sign,
                                float *out)
{
    int P, a;
    if (!row_lane(N, P, a)) return;
    float s = 0.f;
    for (int e = lrow[P]; e < lrow[P + 1]; ++e) {
        const float *w = W + (int64_t)e * 36 + 6 * a;
        const float *v = in + 6 * (int64_t)lcol[e];
        s += w[0] * v[0] + w[1] * v[1] + w[2] * v[2] + w[3] * v[3] + w[4] * v[4] + w[5] * v[5];
    }
    out[6 * (int64_t)P + a] = sign * s;
}

// out_Q = sign sum_{e = (P, Q)} W_e^T in_P   (upper product)
__global__ void lj_upper_kernel(int N, const int *urow, const int *ucol, const int *upos, const float *W,
                                const float *in, float sign, float *out)
{
    int P, a;
    if (!row_lane(N, P, a)) return;
    float s = 0.f;
    for (int e = urow[P]; e < urow[P + 1]; ++e) {
        const float *w = W + (int64_t)upos[e] * 36 + a;
        const float *x = in + 6 * (int64_t)ucol[e];
        s += w[0] * x[0] + w[6] * x[1] + w[12] * x[2] + w[18] * x[3] + w[24] * x[4] + w[30] * x[5];
    }
    out[6 * (int64_t)P + a] = sign * s;
}

__global__ void lj_add_kernel(int64_t n, const float *x, float *y)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) y[i] += x[i];
}

// sum of |x_s|^2 over the samples' full saddle vectors into acc[1]
__global__ void lj_sumsq_kernel(int64_t n, const float *x, double *acc)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = x[i];
        s += v * v;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(acc + 1, red[0]);
}

// loss of one sample: (|g - (b_F - x_F)|^2 + |g - x_U|^2) / sum_s |x_s|^2 with
// g = x_k unpermuted; the gradient with respect to x_k (permuted) into xbar
__global__ void lj_loss_kernel(int N, const int *perm, const float *xk, const float *x17, const float *b17,
                               double *acc, float *xbar)
{
    __shared__ double red[256];
    int P, a;
    const bool ok = row_lane(N, P, a);
    double l = 0.0;
    if (ok) {
        const int64_t i = perm[P], n11 = 11 * (int64_t)N;
        const float g = xk[6 * (int64_t)P + a];
        const float y1 = b17[6 * i + a] - x17[6 * i + a];
        const float y2 = x17[n11 + 6 * i + a];
        const double w = 1.0 / acc[1];
        const float d1 = g - y1, d2 = g - y2;
        l = ((double)d1 * d1 + (double)d2 * d2) * w;
        xbar[6 * (int64_t)P + a] = (float)(2.0 * w) * (d1 + d2);
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(acc, red[0]);
}

// gradient of the pair blocks: gW_e[a][c] = sum_terms coef * A_t[P][a] B_t[Q][c], e = (P, Q)
struct GTerms {
    const float *A[64];
    const float *B[64];
    float coef[64];
    int n;
};
__global__ void lj_gw_kernel(int64_t E, const int *erow, const int *lcol, GTerms T, float *gW)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= E * 6) return;
    const int64_t e = t / 6;
    const int a = (int)(t - e * 6);
    const int64_t P = erow[e], Q = lcol[e];
    float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < T.n; ++k) {
        const float x = T.coef[k] * T.A[k][6 * P + a];
        const float *y = T.B[k] + 6 * Q;
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[c] += x * y[c];
    }
    for (int c = 0; c < 6; ++c) gW[e * 36 + a * 6 + c] = acc[c];
}
// gradient of the inverse pivots: gDinv_P[a][c] = sum_terms coef * A_t[P][a] B_t[P][c]
__global__ void lj_gd_kernel(int N, GTerms T, float *gD)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * 6) return;
    const int64_t P = t / 6;
    const int a = (int)(t - P * 6);
    float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < T.n; ++k) {
        const float x = T.coef[k] * T.A[k][6 * P + a];
        const float *y = T.B[k] + 6 * P;
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[c] += x * y[c];
    }
    for (int c = 0; c < 6; ++c) gD[P * 36 + a * 6 + c] = acc[c];
}

// assembly backward of one pivot pass, per particle: Dbar = -Dinv sym(gDinv) Dinv;
// pivot-shift gradients; per lower pair the fill-weight gradient
// -s_P gamma'(c) <Dbar, A F A^T> with the pivot's level s_P = 2^-level (no output gradient
// at level 4).  dbar_direct: gDinv already holds the gradient with respect to D itself
// (symmetric); accumulate: add to goutn and goute[.., 6] (every pass but the last, which
// runs first); Dbar_out (may be null): Dbar of every particle (zero at level 4)
__global__ void lp_node_D_bwd_kernel(int N, const int *lrow, const int *lcol, const float *Dinv, const float *gDinv,
                                     int dbar_direct, const float *oute, const float *outn, float lambda,
                                     const float *Ablk, const float *Finv, const int *lvl, int accumulate,
                                     float *Dbar_out, float *goutn, float *goute)
{
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= N) return;
    float* go = goutn + (int64_t)P * lp::ON;
    if (!accumulate)
        for (int c = 0; c < lp::ON; ++c) go[c] = 0.f;
    const int level = lvl[P];
    if (level >= LP_LEVELS) {   // D_P = A_PP + lambda S does not depend on the outputs
        if (!accumulate)
            for (int e = lrow[P]; e < lrow[P + 1]; ++e) goute[(int64_t)e * lp::OE + 6] = 0.f;
        if (Dbar_out)
            for (int q = 0; q < 36; ++q) Dbar_out[(int64_t)P * 36 + q] = 0.f;
        return;
    }
    const float sc = ldexpf(1.0f, -level);
    const float *Di = Dinv + (int64_t)P * 36, *G = gDinv + (int64_t)P * 36;
    float S[36], T[36], Db[36];
    for (int a = 0; a < 6; ++a)
        for (int c = 0; c < 6; ++c) S[a * 6 + c] = 0.5f * (G[a * 6 + c] + G[c * 6 + a]);
    if (dbar_direct) {
        for (int q = 0; q < 36; ++q) Db[q] = S[q];
    } else {
        for (int a = 0; a < 6; ++a)
            for (int c = 0; c < 6; ++c) {
                float s = 0.f;
                for (int q = 0; q < 6; ++q) s += Di[a * 6 + q] * S[q * 6 + c];
                T[a * 6 + c] = s;
            }
        for (int a = 0; a < 6; ++a)
            for (int c = 0; c < 6; ++c) {
                float s = 0.f;
                for (int q = 0; q < 6; ++q) s += T[a * 6 + q] * Di[q * 6 + c];
                Db[a * 6 + c] = -s;
            }
    }
    if (Dbar_out)
        for (int q = 0; q < 36; ++q) Dbar_out[(int64_t)P * 36 + q] = Db[q];
    const float g0 = lambda * expf(outn[(int64_t)P * lp::ON]) * (Db[0] + Db[7] + Db[14]);
    const float g1 = lambda * (4.0f / 3.0f) * expf(outn[(int64_t)P * lp::ON + 1]) * (Db[21] + Db[28] + Db[35]);
    go[0] = accumulate ? go[0] + g0 : g0;
    go[1] = accumulate ? go[1] + g1 : g1;
    for (int e = lrow[P]; e < lrow[P + 1]; ++e) {
        const float *Ae = Ablk + (int64_t)e * 36, *Ti = Finv + (int64_t)lcol[e] * 36;
        float AT[36];   // A F
        for (int a = 0; a < 6; ++a)
            for (int c = 0; c < 6; ++c) {
                float s = 0.f;
                for (int q = 0; q < 6; ++q) s += Ae[a * 6 + q] * Ti[q * 6 + c];
                AT[a * 6 + c] = s;
            }
        float gg = 0.f;
        for (int a = 0; a < 6; ++a)
            for (int c = 0; c < 6; ++c) {
                float s = 0.f;
                for (int q = 0; q < 6; ++q) s += AT[a * 6 + q] * Ae[c * 6 + q];
                gg += Db[a * 6 + c] * s;
            }
        const float gc = -sc * gg * lp_gamma_d(oute[(int64_t)e * lp::OE + 6]);
        float *gp = goute + (int64_t)e * lp::OE + 6;
        *gp = accumulate ? *gp + gc : gc;
    }
}

// one-step fill correction of the pair blocks (IC(0)'s W_PQ = A_PQ - sum_R W_PR D_R^-1 W_QR^T
// with the pair blocks for W on the right): W_e -= sum_R A_PR Dinv_R A_QR^T over the common
// lower neighbours R of the edge's particles (R < Q < P; both lower lists are sorted), with
// the factor's final pivots.  One thread per lower block.
__device__ __forceinline__ void lp_mat6_atb(const float *A, const float *M, const float *B, float *out, float w,
                                            bool accumulate)
{
    // out (+)= w * A M B^T
    float T[36];
    for (int a = 0; a < 6; ++a)
        for (int c = 0; c < 6; ++c) {
            float s = 0.f;
            for (int q = 0; q < 6; ++q) s += A[a * 6 + q] * M[q * 6 + c];
            T[a * 6 + c] = s;
        }
    for (int a = 0; a < 6; ++a)
        for (int c = 0; c < 6; ++c) {
            float s = 0.f;
            for (int q = 0; q < 6; ++q) s += T[a * 6 + q] * B[c * 6 + q];
            out[a * 6 + c] = (accumulate ? out[a * 6 + c] : 0.f) + w * s;
        }
}

__global__ void lp_wfix_kernel(int64_t E, const int *erow, const int *lcol, const int *lrow, const float *Ablk,
                               const float *Dinv, float *W)
{
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int P = erow[e], Q = lcol[e];
    int i = lrow[P], j = lrow[Q];
    const int ie = lrow[P + 1], je = lrow[Q + 1];
    float acc[36];
    bool any = false;
    while (i < ie && j < je) {
        const int rp = lcol[i], rq = lcol[j];
        if (rp < rq) { ++i; continue; }
        if (rq < rp) { ++j; continue; }
        lp_mat6_atb(Ablk + (int64_t)i * 36, Dinv + (int64_t)rp * 36, Ablk + (int64_t)j * 36, acc, 1.0f, any);
        any = true;
        ++i; ++j;
    }
    if (!any) return;
    float *w = W + e * 36;
    for (int q = 0; q < 36; ++q) w[q] -= acc[q];
}

// backward of the fill correction: the gradient with respect to the final inverse pivots,
// G_R += -A_PR^T gW_e A_QR for every common lower neighbour R of edge e = (P, Q)
__global__ void lp_wfix_bwd_kernel(int64_t E, const int *erow, const int *lcol, const int *lrow, const float *Ablk,
                                   const float *gW, float *G)
{
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int P = erow[e], Q = lcol[e];
    int i = lrow[P], j = lrow[Q];
    const int ie = lrow[P + 1], je = lrow[Q + 1];
    const float *g = gW + e * 36;
    while (i < ie && j < je) {
        const int rp = lcol[i], rq = lcol[j];
        if (rp < rq) { ++i; continue; }
        if (rq < rp) { ++j; continue; }
        // A_PR^T gW A_QR: (A_PR^T) (gW) (A_QR^T)^T
        const float *Ap = Ablk + (int64_t)i * 36, *Aq = Ablk + (int64_t)j * 36;
        float At[36], Bt[36], out[36];
        for (int a = 0; a < 6; ++a)
            for (int c = 0; c < 6; ++c) {
                At[a * 6 + c] = Ap[c * 6 + a];
                Bt[a * 6 + c] = Aq[c * 6 + a];
            }
        lp_mat6_atb(At, g, Bt, out, -1.0f, false);
        float *gr = G + (int64_t)rp * 36;
        for (int q = 0; q < 36; ++q) atomicAdd(gr + q, out[q]);
        ++i; ++j;
    }
}

// adds the fill correction's gradient G (with respect to the final inverse pivots) to gDinv:
// as it is when gDinv is itself a gradient with respect to the inverse pivots, as
// -Dinv sym(G) Dinv when it holds the gradient with respect to the pivots (dbar_direct)
__global__ void lp_wfix_dbar_kernel(int N, int dbar_direct, const float *Dinv, const float *G, float *gDinv)
{
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= N) return;
    const float *g = G + (int64_t)P * 36;
    float *o = gDinv + (int64_t)P * 36;
    if (!dbar_direct) {
        for (int q = 0; q < 36; ++q) o[q] += g[q];
        return;
    }
    float S[36];
    for (int a = 0; a < 6; ++a)
        for (int c = 0; c < 6; ++c) S[a * 6 + c] = 0.5f * (g[a * 6 + c] + g[c * 6 + a]);
    const float *Di = Dinv + (int64_t)P * 36;
    float out[36];
    lp_mat6_atb(Di, S, Di, out, -1.0f, false);   // Dinv symmetric: Dinv S Dinv^T = Dinv S Dinv
    for (int q = 0; q < 36; ++q) o[q] += out[q];
}

// from a pivot pass to the one before it: the gradient with respect to the earlier pass's
// pivots D1_Q, from the later pass's Dbar2 of every row particle R with a lower block
// (R, Q): the gradient with respect to D1_Q^-1 is X = sum -s_R gamma_e A_e^T Dbar2_R A_e
// (levels of the later pass, none at level 4), and D1bar_Q = -D1_Q^-1 sym(X) D1_Q^-1.
// The column entries of Q are its upper list (ucol: the row particle, upos: the block).
__global__ void lp_fill_bwd_kernel(int N, const int *urow, const int *ucol, const int *upos, const float *Ablk,
                                   const float *oute, const int *lvl2, const float *Dbar2, const float *D1inv,
                                   float *D1bar)
{
    const int Q = blockIdx.x * blockDim.x + threadIdx.x;
    if (Q >= N) return;
    float X[36];
    for (int q = 0; q < 36; ++q) X[q] = 0.f;
    for (int u = urow[Q]; u < urow[Q + 1]; ++u) {
        const int R = ucol[u], e = upos[u];
        if (lvl2[R] >= LP_LEVELS) continue;
        const float w = -ldexpf(1.0f, -lvl2[R]) * lp_gamma(oute[(int64_t)e * lp::OE + 6]);
        const float *A = Ablk + (int64_t)e * 36, *Db = Dbar2 + (int64_t)R * 36;
        float T[36];   // Dbar2_R A
        for (int a = 0; a < 6; ++a)
            for (int c = 0; c < 6; ++c) {
                float s = 0.f;
                for (int q = 0; q < 6; ++q) s += Db[a * 6 + q] * A[q * 6 + c];
                T[a * 6 + c] = s;
            }
        for (int a = 0; a < 6; ++a)
            for (int c = 0; c < 6; ++c) {
                float s = 0.f;
                for (int q = 0; q < 6; ++q) s += A[q * 6 + a] * T[q * 6 + c];
                X[a * 6 + c] += w * s;
            }
    }
    const float *Di = D1inv + (int64_t)Q * 36;
    float S[36], T[36];
    for (int a = 0; a < 6; ++a)
        for (int c = 0; c < 6; ++c) S[a * 6 + c] = 0.5f * (X[a * 6 + c] + X[c * 6 + a]);
    for (int a = 0; a < 6; ++a)
        for (int c = 0; c < 6; ++c) {
            float s = 0.f;
            for (int q = 0; q < 6; ++q) s += Di[a * 6 + q] * S[q * 6 + c];
            T[a * 6 + c] = s;
        }
    for (int a = 0; a < 6; ++a)
        for (int c = 0; c < 6; ++c) {
            float s = 0.f;
            for (int q = 0; q < 6; ++q) s += T[a * 6 + q] * Di[q * 6 + c];
            D1bar[(int64_t)Q * 36 + a * 6 + c] = -s;
        }
}

// assembly backward, per edge: multiplier gradients (1 - tanh^2 c_k) <gW, basis_k>; other outputs 0
__global__ void lp_edge_W_bwd_kernel(int64_t E, const float *geo, const float *oute, const float *gW, float *goute)
{
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= E) return;
    const float *g = geo + e * 8;
    const float r[3] = {g[0], g[1], g[2]};
    const float ph[5] = {g[3], g[4], g[5], g[6], g[7]};
    float E3[3][3];
    eps_r(r, E3);
    const float *gw = gW + e * 36;
    float *o = goute + e * lp::OE;
    for (int k = 0; k < 6; ++k) {
        float s = 0.f;
        for (int a = 0; a < 6; ++a)
            for (int c = 0; c < 6; ++c) s += gw[a * 6 + c] * basis_entry(k, a, c, r, E3, ph);
        const float th = tanhf(oute[e * lp::OE + k]);
        o[k] = (1.0f - th * th) * s;
    }
    for (int k = 7; k < lp::OE; ++k) o[k] = 0.f;
}

// gz = g * silu'(z) (act) or g (linear); FP32 and BF16 copies
__global__ void lp_act_bwd_kernel(int64_t n, const float *g, const float *z, float *gz, uint16_t *gbf)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = z ? g[i] * silu_d(z[i]) : g[i];
    gz[i] = v;
    gbf[i] = f2bf(v);
}

// particle gradients from the message inputs: gh_P (+)= sum over lower edges of
// gx[e][0:64] + sum over upper edges of gx[e][64:128]
__global__ void lp_scatter_h_kernel(int N, const int *lrow, const int *urow, const int *upos, const float *gx,
                                    const float *base, float *gh)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * lp::H) return;
    const int P = (int)(t / lp::H), c = (int)(t % lp::H);
    float s = base ? base[t] : 0.f;
    for (int e = lrow[P]; e < lrow[P + 1]; ++e) s += gx[(int64_t)e * 3 * lp::H + c];
    for (int u = urow[P]; u < urow[P + 1]; ++u) s += gx[(int64_t)upos[u] * 3 * lp::H + lp::H + c];
    gh[t] = s;
}

// ge (+)= gx[e][128:192]
__global__ void lp_edge_resid_kernel(int64_t E, const float *gx, const float *base, float *ge)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= E * lp::H) return;
    const int64_t e = t / lp::H;
    const int c = (int)(t - e * lp::H);
    ge[t] = (base ? base[t] : 0.f) + gx[e * 3 * lp::H + 2 * lp::H + c];
}

// message gradient: gm = ge + gagg_P + gagg_Q with gagg = gxn[:, 64:128]; gh2 = gh + gxn[:, 0:64]
__global__ void lp_gm_kernel(int64_t E, const int *erow, const int *lcol, const float *ge, const float *gxn, float *gm)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= E * lp::H) return;
    const int64_t e = t / lp::H;
    const int c = (int)(t - e * lp::H);
    gm[t] = ge[t] + gxn[(int64_t)erow[e] * 2 * lp::H + lp::H + c] + gxn[(int64_t)lcol[e] * 2 * lp::H + lp::H + c];
}
__global__ void lp_gh_upd_kernel(int N, const float *gh, const float *gxn, float *gh2)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * lp::H) return;
    const int64_t P = t / lp::H;
    const int c = (int)(t - P * lp::H);
    gh2[t] = gh[t] + gxn[P * 2 * lp::H + c];
}

// Adam (Kingma & Ba 2015) with bias correction
__global__ void lp_adam_kernel(int64_t n, float *w, const float *g, float *m, float *v, float lr, float b1, float b2,
                               float eps, float c1, float c2)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 copies of one layer: wt[n][k] and wn[k][n] from w[k][n]
__global__ void lp_pack_kernel(int ki, int no, const float *w, uint16_t *wt, uint16_t *wn)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ki * no) return;
    const int k = q / no, n = q % no;
    const uint16_t b = f2bf(w[q]);
    wn[q] = b;
    wt[n * ki + k] = b;
}

// ------------------------------------------------- supervised factor targets
// IC(0)'s own block LDL^T: pivots D_P = L_PP L_PP^T, pair blocks W_PQ = L_PQ L_QQ^T.
// loss += |D_P - D_P^IC|^2 / (N |D_P^IC|^2) and gDbar = its gradient with respect to D_P
__global__ void lp_pivot_target_kernel(int N, const float *Ldiag, const float *D, double *loss, float *gD)
{
    __shared__ double red[128];
    const int P = blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (P < N) {
        const float *Lp = Ldiag + (int64_t)P * 36, *Dp = D + (int64_t)P * 36;
        float T[36];
        float nt = 0.f;
        for (int a = 0; a < 6; ++a)
            for (int c = 0; c < 6; ++c) {
                float v = 0.f;
                for (int q = 0; q <= (a < c ? a : c); ++q) v += Lp[a * 6 + q] * Lp[c * 6 + q];
                T[a * 6 + c] = v;
                nt += v * v;
            }
        const float w = 1.0f / ((float)N * fmaxf(nt, 1e-30f));
        for (int q = 0; q < 36; ++q) {
            const float d = Dp[q] - T[q];
            l += (double)w * d * d;
            gD[(int64_t)P * 36 + q] = 2.0f * w * d;
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 64; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// loss += |W_e - W_e^IC|^2 / (E |A_e|^2) and gW = its gradient, e = (P, Q)
__global__ void lp_pair_target_kernel(int64_t E, const int *lcol, const float *Lblk, const float *Ldiag,
                                      const float *Ablk, const float *W, double *loss, float *gW)
{
    __shared__ double red[128];
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    double l = 0.0;
    if (e < E) {
        const float *Le = Lblk + e * 36, *Lq = Ldiag + (int64_t)lcol[e] * 36, *Ae = Ablk + e * 36, *We = W + e * 36;
        float na = 0.f;
        for (int q = 0; q < 36; ++q) na += Ae[q] * Ae[q];
        const float w = 1.0f / ((float)E * fmaxf(na, 1e-30f));
        for (int a = 0; a < 6; ++a)
            for (int c = 0; c < 6; ++c) {
                float t = 0.f;   // (L_PQ L_QQ^T)_ac, L_QQ lower triangular
                for (int q = 0; q <= c; ++q) t += Le[a * 6 + q] * Lq[c * 6 + q];
                const float d = We[a * 6 + c] - t;
                l += (double)w * d * d;
                gW[e * 36 + a * 6 + c] = 2.0f * w * d;
            }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 64; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// =================================================================== host side
namespace {
template <typename T>
int lalloc(fsd_ctx *ctx, T *&p, int64_t n)
{
    if (p) hipFree(p);
    p = nullptr;
    FSD_HIP(hipMalloc(&p, sizeof(T) * (size_t)(n > 0 ? n : 1)));
    return 0;
}
#define LRC(call)                      \
    do {                               \
        const int r_ = (call);         \
        if (r_) return r_;             \
    } while (0)

template <int KI, int NO, int ACT>
void mm(fsd_ctx *ctx, int64_t R, const uint16_t *X, const uint16_t *Bt, const float *bias, const float *res, float *Z,
        float *A, float *Y)
{
    if (R <= 0) return;
    hipLaunchKernelGGL((lp_mm_kernel<KI, NO, ACT>), dim3(nblk(R, 64)), dim3(256), 0, ctx->stream, (int)R, X, Bt, bias,
                       res, Z, A, Y);
}

// layer l forward: X [R][ki] -> outputs
void layer_fwd(fsd_ctx *ctx, int l, int64_t R, const uint16_t *X, const float *res, float *Z, float *A, float *Y)
{
    LearnedState &s = *ctx->lp;
    const lp::Layer &L = s.L[l];
    const uint16_t *Bt = s.d_wt + L.bf_off;
    const float *b = s.d_w + L.b_off;
    if (L.ki == lp::FIN && L.no == lp::H) mm<lp::FIN, lp::H, 1>(ctx, R, X, Bt, b, res, Z, A, Y);
    else if (L.ki == 3 * lp::H && L.no == lp::H) mm<3 * lp::H, lp::H, 1>(ctx, R, X, Bt, b, res, Z, A, Y);
    else if (L.ki == 2 * lp::H && L.no == lp::H) mm<2 * lp::H, lp::H, 1>(ctx, R, X, Bt, b, res, Z, A, Y);
    else if (L.ki == 3 * lp::H && L.no == lp::OE) mm<3 * lp::H, lp::OE, 0>(ctx, R, X, Bt, b, res, Z, A, Y);
    else mm<lp::H, lp::ON, 0>(ctx, R, X, Bt, b, res, Z, A, Y);
}

// layer l backward: given gz (FP32) and its BF16 copy for R rows with inputs X,
// writes the weight and bias gradients and (gx != null) gx = gz W^T [R][ki]
int layer_bwd(fsd_ctx *ctx, int l, int64_t R, const uint16_t *X, const float *gz, const uint16_t *gzbf, float *gx)
{
    LearnedState &s = *ctx->lp;
    const lp::Layer &L = s.L[l];
    const int nchunk = (int)std::max<int64_t>(1, (R + 255) / 256);
    const int64_t need = (int64_t)nchunk * (L.ki * L.no + L.no);
    if (need > s.part_cap) {
        FSD_HIP(hipStreamSynchronize(ctx->stream));
        LRC(lalloc(ctx, s.part, need + need / 2));
        s.part_cap = need + need / 2;
    }
    const dim3 g(nchunk), b(256);
    if (L.ki == lp::FIN && L.no == lp::H) hipLaunchKernelGGL((lp_dw_kernel<lp::FIN, lp::H>), g, b, 0, ctx->stream, (int)R, X, gz, s.part);
    else if (L.ki == 3 * lp::H && L.no == lp::H) hipLaunchKernelGGL((lp_dw_kernel<3 * lp::H, lp::H>), g, b, 0, ctx->stream, (int)R, X, gz, s.part);
    else if (L.ki == 2 * lp::H && L.no == lp::H) hipLaunchKernelGGL((lp_dw_kernel<2 * lp::H, lp::H>), g, b, 0, ctx->stream, (int)R, X, gz, s.part);
    else if (L.ki == 3 * lp::H && L.no == lp::OE) hipLaunchKernelGGL((lp_dw_kernel<3 * lp::H, lp::OE>), g, b, 0, ctx->stream, (int)R, X, gz, s.part);
    else hipLaunchKernelGGL((lp_dw_kernel<lp::H, lp::ON>), g, b, 0, ctx->stream, (int)R, X, gz, s.part);
    const int kn = L.ki * L.no;
    hipLaunchKernelGGL(lp_dw_reduce_kernel, dim3(nblk(kn + L.no, 256)), dim3(256), 0, ctx->stream, nchunk, kn, L.no,
                       s.part, s.d_grad + L.w_off, s.d_grad + L.b_off);
    if (gx && R > 0) {
        const uint16_t *Bt = s.d_wn + L.bf_off + (int64_t)L.ki * L.no;   // natural [ki][no] = Bt of gz W^T
        if (L.no == lp::H && L.ki == lp::FIN) mm<lp::H, lp::FIN, 0>(ctx, R, gzbf, Bt, nullptr, nullptr, nullptr, nullptr, gx);
        else if (L.no == lp::H && L.ki == 3 * lp::H) mm<lp::H, 3 * lp::H, 0>(ctx, R, gzbf, Bt, nullptr, nullptr, nullptr, nullptr, gx);
        else if (L.no == lp::H && L.ki == 2 * lp::H) mm<lp::H, 2 * lp::H, 0>(ctx, R, gzbf, Bt, nullptr, nullptr, nullptr, nullptr, gx);
        else if (L.no == lp::OE && L.ki == 3 * lp::H) mm<lp::OE, 3 * lp::H, 0>(ctx, R, gzbf, Bt, nullptr, nullptr, nullptr, nullptr, gx);
        else mm<lp::ON, lp::H, 0>(ctx, R, gzbf, Bt, nullptr, nullptr, nullptr, nullptr, gx);
    }
    FSD_HIP(hipGetLastError());
    return 0;
}

int pack_weights(fsd_ctx *ctx)
{
    LearnedState &s = *ctx->lp;
    for (int l = 0; l < lp::NLAYER; ++l) {
        const lp::Layer &L = s.L[l];
        const int kn = L.ki * L.no;
        hipLaunchKernelGGL(lp_pack_kernel, dim3(nblk(kn, 256)), dim3(256), 0, ctx->stream, L.ki, L.no,
                           s.d_w + L.w_off, s.d_wt + L.bf_off, s.d_wn + L.bf_off + kn);
    }
    FSD_HIP(hipGetLastError());
    return 0;
}

// inverse pivots that pass p's fill terms use (the uncorrected ones for the first pass) and
// that pass p writes (the factor's own for the last)
const float *pass_finv(const LearnedState &s, int64_t N, int p)
{
    return p == 0 ? s.Dtinv : s.Dpinv + (int64_t)(p - 1) * N * 36;
}
float *pass_dinv(const LearnedState &s, int64_t N, int p)
{
    return p == lp::NPASS - 1 ? s.Dinv : s.Dpinv + (int64_t)p * N * 36;
}

int grow_edges(fsd_ctx *ctx, int64_t E)
{
    LearnedState &s = *ctx->lp;
    if (s.ecap > 0 && E <= s.ecap) return 0;
    FSD_HIP(hipStreamSynchronize(ctx->stream));
    const int64_t c = E + E / 4 + 64;
    LRC(lalloc(ctx, s.xe0, c * lp::FIN));
    LRC(lalloc(ctx, s.ze0, c * lp::H));
    for (int l = 0; l <= lp::NMP; ++l) {
        LRC(lalloc(ctx, s.e[l], c * lp::H));
        LRC(lalloc(ctx, s.xm[l], c * 3 * lp::H));
    }
    for (int l = 0; l < lp::NMP; ++l) {
        LRC(lalloc(ctx, s.zm[l], c * lp::H));
        LRC(lalloc(ctx, s.m[l], c * lp::H));
    }
    LRC(lalloc(ctx, s.oute, c * lp::OE));
    LRC(lalloc(ctx, s.geo, c * 8));
    LRC(lalloc(ctx, s.W, c * 36));
    LRC(lalloc(ctx, s.erow, c));
    // backward (edge side)
    LRC(lalloc(ctx, s.ge, c * lp::H));
    LRC(lalloc(ctx, s.gm, c * lp::H));
    LRC(lalloc(ctx, s.gze, c * lp::H));
    LRC(lalloc(ctx, s.gxe, c * 3 * lp::H));
    LRC(lalloc(ctx, s.goute, c * lp::OE));
    LRC(lalloc(ctx, s.gW, c * 36));
    const int64_t gb = std::max<int64_t>(c, ctx->N) * lp::H;
    LRC(lalloc(ctx, s.gbf, gb));
    s.ecap = c;
    return 0;
}
}  // namespace

int lp_init(fsd_ctx *ctx)
{
    if (ctx->lp) return 0;
    ctx->lp = new LearnedState();
    LearnedState &s = *ctx->lp;
    lp::layers(s.L, &s.nparam, &s.nbf);
    const int64_t N = ctx->N;
    LRC(lalloc(ctx, s.d_w, s.nparam));
    LRC(lalloc(ctx, s.d_grad, s.nparam));
    LRC(lalloc(ctx, s.d_m, s.nparam));
    LRC(lalloc(ctx, s.d_v, s.nparam));
    // one buffer of BF16 copies, per layer [transposed ki*no | natural ki*no]
    LRC(lalloc(ctx, s.d_wn, s.nbf));
    s.d_wt = s.d_wn;
    FSD_HIP(hipMemsetAsync(s.d_w, 0, sizeof(float) * s.nparam, ctx->stream));
    FSD_HIP(hipMemsetAsync(s.d_m, 0, sizeof(float) * s.nparam, ctx->stream));
    FSD_HIP(hipMemsetAsync(s.d_v, 0, sizeof(float) * s.nparam, ctx->stream));
    LRC(lalloc(ctx, s.xn0, N * lp::FIN));
    LRC(lalloc(ctx, s.zn0, N * lp::H));
    for (int l = 0; l <= lp::NMP; ++l) LRC(lalloc(ctx, s.h[l], N * lp::H));
    for (int l = 0; l < lp::NMP; ++l) {
        LRC(lalloc(ctx, s.xu[l], N * 2 * lp::H));
        LRC(lalloc(ctx, s.zu[l], N * lp::H));
    }
    LRC(lalloc(ctx, s.agg, N * lp::H));
    LRC(lalloc(ctx, s.xh, N * lp::H));
    LRC(lalloc(ctx, s.outn, N * lp::ON));
    LRC(lalloc(ctx, s.Dinv, N * 36));
    LRC(lalloc(ctx, s.D, N * 36));
    LRC(lalloc(ctx, s.Dtinv, N * 36));
    LRC(lalloc(ctx, s.gh, N * lp::H));
    LRC(lalloc(ctx, s.gh2, N * lp::H));
    LRC(lalloc(ctx, s.gzn, N * lp::H));
    LRC(lalloc(ctx, s.gxn, N * 2 * lp::H));
    LRC(lalloc(ctx, s.goutn, N * lp::ON));
    LRC(lalloc(ctx, s.gDinv, N * 36));
    LRC(lalloc(ctx, s.d_fail, 4));
    LRC(lalloc(ctx, s.lvlp, N * lp::NPASS));
    LRC(lalloc(ctx, s.Dpinv, N * 36 * (lp::NPASS - 1)));
    LRC(lalloc(ctx, s.Dbar2, N * 36));
    LRC(lalloc(ctx, s.d_loss, 4));
    FSD_HIP(hipHostMalloc((void **)&s.h_fail, sizeof(int) * 4, hipHostMallocDefault));
    FSD_HIP(hipHostMalloc((void **)&s.h_loss, sizeof(double) * 4, hipHostMallocDefault));
    LRC(pack_weights(ctx));
    ctx->lp->ncap = ctx->N;
    FSD_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

void lp_free(fsd_ctx *ctx)
{
    if (!ctx->lp) return;
    LearnedState &s = *ctx->lp;
    void *ptrs[] = {s.d_w, s.d_grad, s.d_m, s.d_v, s.d_wn, s.xn0, s.zn0, s.agg, s.xh, s.outn, s.Dinv, s.Dtinv, s.D,
                    s.xe0, s.ze0, s.oute, s.geo, s.W, s.erow, s.d_fail, s.gh, s.gh2, s.ge, s.gm, s.gzn, s.gze, s.gxe,
                    s.gxn, s.goute, s.goutn, s.gW, s.gDinv, s.gbf, s.part, s.jv, s.d_loss, s.d_wbak, s.lvlp, s.Dpinv, s.Dbar2};
    for (void *p : ptrs)
        if (p) hipFree(p);
    for (int l = 0; l <= lp::NMP; ++l) {
        if (s.h[l]) hipFree(s.h[l]);
        if (s.e[l]) hipFree(s.e[l]);
        if (s.xm[l]) hipFree(s.xm[l]);
    }
    for (int l = 0; l < lp::NMP; ++l) {
        if (s.xu[l]) hipFree(s.xu[l]);
        if (s.zu[l]) hipFree(s.zu[l]);
        if (s.zm[l]) hipFree(s.zm[l]);
        if (s.m[l]) hipFree(s.m[l]);
    }
    if (s.h_fail) hipHostFree(s.h_fail);
    if (s.h_loss) hipHostFree(s.h_loss);
    delete ctx->lp;
    ctx->lp = nullptr;
}

int64_t lp_nparam(fsd_ctx *ctx) { return ctx->lp ? ctx->lp->nparam : 0; }
bool lp_active(const fsd_ctx *ctx) { return ctx->lp && ctx->lp->active; }
bool lp_have_weights(const fsd_ctx *ctx) { return ctx->lp && ctx->lp->have_w; }
void lp_last_factor(const fsd_ctx *ctx, float *lambda, int *shrunk, int *plain)
{
    if (!ctx->lp) return;
    *lambda = ctx->lp->lambda;
    *shrunk = ctx->lp->shrunk;
    *plain = ctx->lp->plain;
}
void lp_set_active(fsd_ctx *ctx, bool on)
{
    if (ctx->lp) ctx->lp->active = on;
}

int lp_set_weights(fsd_ctx *ctx, const float *h_w, int64_t n)
{
    LRC(lp_init(ctx));
    LearnedState &s = *ctx->lp;
    if (n != s.nparam) {
        ctx->err = "learned weights: expected " + std::to_string(s.nparam) + " values";
        return FSD_E_PARAM;
    }
    FSD_HIP(hipMemcpyAsync(s.d_w, h_w, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    FSD_HIP(hipMemsetAsync(s.d_m, 0, sizeof(float) * n, ctx->stream));
    FSD_HIP(hipMemsetAsync(s.d_v, 0, sizeof(float) * n, ctx->stream));
    LRC(pack_weights(ctx));
    FSD_HIP(hipStreamSynchronize(ctx->stream));
    s.have_w = true;
    return 0;
}

// Adam moments (m, v) to or from host buffers (either may be null)
int lp_adam_state(fsd_ctx *ctx, float *h_m, float *h_v, int64_t n, bool set)
{
    LRC(lp_init(ctx));
    LearnedState &s = *ctx->lp;
    if (n != s.nparam) {
        ctx->err = "learned optimiser state: expected " + std::to_string(s.nparam) + " values";
        return FSD_E_PARAM;
    }
    const hipMemcpyKind kind = set ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    if (h_m) FSD_HIP(hipMemcpyAsync(set ? (void *)s.d_m : (void *)h_m, set ? (const void *)h_m : (const void *)s.d_m,
                                    sizeof(float) * n, kind, ctx->stream));
    if (h_v) FSD_HIP(hipMemcpyAsync(set ? (void *)s.d_v : (void *)h_v, set ? (const void *)h_v : (const void *)s.d_v,
                                    sizeof(float) * n, kind, ctx->stream));
    FSD_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int lp_get_weights(fsd_ctx *ctx, float *h_w, int64_t n, bool grad)
{
    LRC(lp_init(ctx));
    LearnedState &s = *ctx->lp;
    if (n != s.nparam) {
        ctx->err = "learned weights: expected " + std::to_string(s.nparam) + " values";
        return FSD_E_PARAM;
    }
    FSD_HIP(hipMemcpyAsync(h_w, grad ? s.d_grad : s.d_w, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
    FSD_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

// Features, network forward and factor assembly on the current pattern (the
// IC(0) path's prune, RCM, lower pattern and pair blocks must be built).  The
// pivot shift lambda starts at 1 and doubles while an uncorrected pivot
// A_PP + lambda S is not positive definite (the relaxer rule of IC(0),
// Precondition.cu:1274-1288, applied to the factor's own pivots).
// *failed: bit 0 non-finite network outputs, bit 1 no lambda up to 1e7.
int lp_forward(fsd_ctx *ctx, const float4 *pos, bool train, int inject, int *failed)
{
    LRC(lp_init(ctx));
    LearnedState &s = *ctx->lp;
    const int N = ctx->N;
    const int64_t E = ctx->pch.nlow;
    LRC(grow_edges(ctx, E));
    LubTable t{ctx->d_res_dist, ctx->d_res_vals, ctx->table_min, ctx->table_dr};
    hipEvent_t pev;
    prof_begin(ctx, PROF_NET, &pev);
    hipLaunchKernelGGL(lp_erow_kernel, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, N, ctx->d_lrow, s.erow);
    if (E)
        hipLaunchKernelGGL(lp_edge_feat_kernel, dim3(nblk(E, 256)), dim3(256), 0, ctx->stream, E, s.erow, ctx->d_lcol,
                           ctx->d_perm, pos, ctx->box, t, s.geo, s.xe0);
    float lambda = 1.0f;
    for (;;) {
        hipLaunchKernelGGL(lp_fill_int_kernel, dim3(1), dim3(64), 0, ctx->stream, 4, 0, s.d_fail);
        hipLaunchKernelGGL(lp_node_feat_kernel, dim3(nblk(N, 128)), dim3(128), 0, ctx->stream, N, ctx->d_lrow,
                           ctx->d_urow, ctx->d_Adiag, lambda, s.xn0, s.Dtinv, s.d_fail);
        FSD_HIP(hipMemcpyAsync(s.h_fail, s.d_fail, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        FSD_HIP(hipStreamSynchronize(ctx->stream));
        if (!(s.h_fail[0] & 2)) break;
        lambda *= 2.0f;
        if (lambda > 1.0e7f) {
            prof_end(ctx, PROF_NET, pev);
            *failed = 2;
            return 0;
        }
    }
    s.lambda = lambda;
    // encoders
    layer_fwd(ctx, lp::L_ENC_N, N, s.xn0, nullptr, train ? s.zn0 : nullptr, nullptr, s.h[0]);
    layer_fwd(ctx, lp::L_ENC_E, E, s.xe0, nullptr, train ? s.ze0 : nullptr, nullptr, s.e[0]);
    for (int l = 0; l < lp::NMP; ++l) {
        if (E)
            hipLaunchKernelGGL(lp_gather_msg_kernel, dim3(nblk(E * 3 * lp::H, 256)), dim3(256), 0, ctx->stream, E,
                               s.erow, ctx->d_lcol, s.h[l], s.e[l], s.xm[l]);
        layer_fwd(ctx, lp::L_MSG + 2 * l, E, s.xm[l], s.e[l], train ? s.zm[l] : nullptr, s.m[l], s.e[l + 1]);
        hipLaunchKernelGGL(lp_agg_kernel, dim3(nblk((int64_t)N * lp::H, 256)), dim3(256), 0, ctx->stream, N,
                           ctx->d_lrow, ctx->d_urow, ctx->d_upos, s.m[l], s.h[l], s.agg, s.xu[l]);
        layer_fwd(ctx, lp::L_UPD + 2 * l, N, s.xu[l], s.h[l], train ? s.zu[l] : nullptr, nullptr, s.h[l + 1]);
    }
    if (E)
        hipLaunchKernelGGL(lp_gather_msg_kernel, dim3(nblk(E * 3 * lp::H, 256)), dim3(256), 0, ctx->stream, E, s.erow,
                           ctx->d_lcol, s.h[lp::NMP], s.e[lp::NMP], s.xm[lp::NMP]);
    layer_fwd(ctx, lp::L_HEAD_E, E, s.xm[lp::NMP], nullptr, nullptr, nullptr, s.oute);
    hipLaunchKernelGGL(lp_to_bf_kernel, dim3(nblk((int64_t)N * lp::H, 256)), dim3(256), 0, ctx->stream,
                       (int64_t)N * lp::H, s.h[lp::NMP], s.xh);
    layer_fwd(ctx, lp::L_HEAD_N, N, s.xh, nullptr, nullptr, nullptr, s.outn);
    if (inject >= 0 && inject < N) {   // failure-path tests: a pivot shift that makes the pivot NaN (not positive)
        const float bad = __builtin_nanf("");
        FSD_HIP(hipMemcpyAsync(s.outn + (int64_t)inject * lp::ON, &bad, sizeof(float), hipMemcpyHostToDevice,
                               ctx->stream));
        FSD_HIP(hipStreamSynchronize(ctx->stream));
    }
    // factor
    if (E)
        hipLaunchKernelGGL(lp_edge_W_kernel, dim3(nblk(E * 36, 256)), dim3(256), 0, ctx->stream, E, s.geo, s.oute, s.W);
    // NPASS parallel passes of the pivot recurrence: the first with the uncorrected neighbour
    // pivots in the fill terms, each later one with the previous pass's (only the last pass's
    // safeguard counts are kept)
    for (int p = 0; p < lp::NPASS; ++p) {
        const bool last = p == lp::NPASS - 1;
        if (last) hipLaunchKernelGGL(lp_fill_int_kernel, dim3(1), dim3(64), 0, ctx->stream, 2, 0, s.d_fail + 1);
        hipLaunchKernelGGL(lp_node_D_kernel, dim3(nblk(N, 64)), dim3(64), 0, ctx->stream, N, ctx->d_lrow,
                           ctx->d_lcol, ctx->d_Adiag, lambda, s.outn, s.oute, ctx->d_Ablk, s.Dtinv, pass_finv(s, N, p),
                           pass_dinv(s, N, p), last ? s.D : (float *)nullptr, s.lvlp + (int64_t)p * N, s.d_fail);
    }
    if (E)   // the pair blocks' one-step fill correction with the final pivots
        hipLaunchKernelGGL(lp_wfix_kernel, dim3(nblk(E, 128)), dim3(128), 0, ctx->stream, E, s.erow, ctx->d_lcol,
                           ctx->d_lrow, ctx->d_Ablk, s.Dinv, s.W);
    FSD_HIP(hipGetLastError());
    prof_end(ctx, PROF_NET, pev);
    FSD_HIP(hipMemcpyAsync(s.h_fail, s.d_fail, sizeof(int) * 3, hipMemcpyDeviceToHost, ctx->stream));
    FSD_HIP(hipStreamSynchronize(ctx->stream));
    *failed = s.h_fail[0];
    s.shrunk = s.h_fail[1];
    s.plain = s.h_fail[2];
    return 0;
}

// G applied by k block Jacobi iterations each way; mode 0: y = G b0 (6N), mode 1:
// the saddle preconditioner of b0 (17N) with G (b_F - b_U)
int lp_apply(fsd_ctx *ctx, const float *b0, const float *b1, const float *bsaddle, float *x, int mode)
{
    LearnedState &s = *ctx->lp;
    const int N = ctx->N, k = ctx->jac_k;
    hipEvent_t ev;
    prof_begin(ctx, PROF_TRSV, &ev);
    const dim3 g(row_blocks(N)), b(256);
    float *t = ctx->d_jac, *v[2] = {ctx->d_jac + 6 * (int64_t)N, ctx->d_jac + 12 * (int64_t)N};
    float *xb[2] = {ctx->d_jac + 18 * (int64_t)N, ctx->d_jac + 12 * (int64_t)N};
    hipLaunchKernelGGL(lj_pre_kernel, g, b, 0, ctx->stream, N, ctx->d_perm, s.Dinv, b0, b1, t, v[0]);
    int cur = 0;
    hipEvent_t e1;
    prof_begin(ctx, PROF_LJIT, &e1);   // the k forward iterations, one event pair
    for (int it = 0; it < k; ++it) {
        hipLaunchKernelGGL(lj_fwd_kernel, g, b, 0, ctx->stream, N, ctx->d_lrow, ctx->d_lcol, s.W, s.Dinv, t, v[cur],
                           (float *)nullptr, v[cur ^ 1]);
        cur ^= 1;
    }
    prof_end_n(ctx, PROF_LJIT, e1, k);
    const float *vk = v[cur];
    // backward iterates alternate between the two buffers vk does not occupy
    float *bo[2] = {xb[0], (vk == v[0]) ? v[1] : v[0]};
    const float *xc = vk;
    for (int it = 0; it < k; ++it) {   // the last iteration writes the preconditioner's output as well
        float *out = bo[it & 1];
        const bool last = it == k - 1;
        hipLaunchKernelGGL(lj_bwd_kernel, g, b, 0, ctx->stream, N, ctx->d_urow, ctx->d_ucol, ctx->d_upos, s.W, s.Dinv,
                           vk, xc, (float *)nullptr, out, last ? ctx->d_perm : (const int *)nullptr,
                           last ? bsaddle : (const float *)nullptr, last ? x : (float *)nullptr, last ? mode : -1);
        xc = out;
    }
    if (k == 0) hipLaunchKernelGGL(lj_post_kernel, g, b, 0, ctx->stream, N, ctx->d_perm, xc, bsaddle, x, mode);
    FSD_HIP(hipGetLastError());
    prof_end(ctx, PROF_TRSV, ev);
    return 0;
}

// Backward of the factor assembly and the network: from gW (pair blocks) and gDinv
// (inverse pivots, or the pivots themselves when dbar_direct) to every parameter's gradient
static int lp_backward(fsd_ctx *ctx, bool dbar_direct)
{
    LearnedState &s = *ctx->lp;
    const int N = ctx->N;
    const int64_t E = ctx->pch.nlow;
    // the pair blocks' fill correction depends on the final pivots: its gradient joins gDinv
    if (E) {
        hipLaunchKernelGGL(lp_fill_kernel, dim3(nblk((int64_t)N * 36, 256)), dim3(256), 0, ctx->stream,
                           (int64_t)N * 36, 0.0f, s.Dbar2);
        hipLaunchKernelGGL(lp_wfix_bwd_kernel, dim3(nblk(E, 128)), dim3(128), 0, ctx->stream, E, s.erow, ctx->d_lcol,
                           ctx->d_lrow, ctx->d_Ablk, s.gW, s.Dbar2);
        hipLaunchKernelGGL(lp_wfix_dbar_kernel, dim3(nblk(N, 128)), dim3(128), 0, ctx->stream, N,
                           dbar_direct ? 1 : 0, s.Dinv, s.Dbar2, s.gDinv);
    }
    // assembly backward: the pivot passes from the last to the first; between two passes, the
    // later pass's dependence on the earlier one's pivots (gradient into gDinv, which the
    // earlier pass then reads as its pivot gradient)
    for (int p = lp::NPASS - 1; p >= 0; --p) {
        const bool last = p == lp::NPASS - 1;
        const int *lvl = s.lvlp + (int64_t)p * N;
        hipLaunchKernelGGL(lp_node_D_bwd_kernel, dim3(nblk(N, 64)), dim3(64), 0, ctx->stream, N, ctx->d_lrow,
                           ctx->d_lcol, pass_dinv(s, N, p), s.gDinv, (last && !dbar_direct) ? 0 : 1, s.oute, s.outn,
                           s.lambda, ctx->d_Ablk, pass_finv(s, N, p), lvl, last ? 0 : 1,
                           p > 0 ? s.Dbar2 : (float *)nullptr, s.goutn, s.goute);
        if (p > 0)
            hipLaunchKernelGGL(lp_fill_bwd_kernel, dim3(nblk(N, 64)), dim3(64), 0, ctx->stream, N, ctx->d_urow,
                               ctx->d_ucol, ctx->d_upos, ctx->d_Ablk, s.oute, lvl, s.Dbar2, pass_finv(s, N, p), s.gDinv);
    }
    if (E) hipLaunchKernelGGL(lp_edge_W_bwd_kernel, dim3(nblk(E, 128)), dim3(128), 0, ctx->stream, E, s.geo, s.oute,
                              s.gW, s.goute);
    // heads
    const int64_t NH = (int64_t)N * lp::H, EH = E * lp::H;
    hipLaunchKernelGGL(lp_act_bwd_kernel, dim3(nblk((int64_t)N * lp::ON, 256)), dim3(256), 0, ctx->stream,
                       (int64_t)N * lp::ON, s.goutn, (const float *)nullptr, s.gzn, s.gbf);
    LRC(layer_bwd(ctx, lp::L_HEAD_N, N, s.xh, s.gzn, s.gbf, s.gh));
    if (E) {
        hipLaunchKernelGGL(lp_act_bwd_kernel, dim3(nblk(E * lp::OE, 256)), dim3(256), 0, ctx->stream, E * lp::OE,
                           s.goute, (const float *)nullptr, s.gze, s.gbf);
    }
    LRC(layer_bwd(ctx, lp::L_HEAD_E, E, s.xm[lp::NMP], s.gze, s.gbf, s.gxe));
    if (E) {
        hipLaunchKernelGGL(lp_scatter_h_kernel, dim3(nblk(NH, 256)), dim3(256), 0, ctx->stream, N, ctx->d_lrow,
                           ctx->d_urow, ctx->d_upos, s.gxe, s.gh, s.gh2);
        std::swap(s.gh, s.gh2);
        hipLaunchKernelGGL(lp_edge_resid_kernel, dim3(nblk(EH, 256)), dim3(256), 0, ctx->stream, E, s.gxe,
                           (const float *)nullptr, s.ge);
    }
    for (int l = lp::NMP - 1; l >= 0; --l) {
        // update: h_l+1 = h_l + silu([h_l | agg_l] Wu + b)
        hipLaunchKernelGGL(lp_act_bwd_kernel, dim3(nblk(NH, 256)), dim3(256), 0, ctx->stream, NH, s.gh, s.zu[l],
                           s.gzn, s.gbf);
        LRC(layer_bwd(ctx, lp::L_UPD + 2 * l, N, s.xu[l], s.gzn, s.gbf, s.gxn));
        hipLaunchKernelGGL(lp_gh_upd_kernel, dim3(nblk(NH, 256)), dim3(256), 0, ctx->stream, N, s.gh, s.gxn, s.gh2);
        std::swap(s.gh, s.gh2);
        if (!E) continue;
        // messages: m_l at both end particles' aggregates, and e_l+1 = e_l + m_l
        hipLaunchKernelGGL(lp_gm_kernel, dim3(nblk(EH, 256)), dim3(256), 0, ctx->stream, E, s.erow, ctx->d_lcol,
                           s.ge, s.gxn, s.gm);
        hipLaunchKernelGGL(lp_act_bwd_kernel, dim3(nblk(EH, 256)), dim3(256), 0, ctx->stream, EH, s.gm, s.zm[l],
                           s.gze, s.gbf);
        LRC(layer_bwd(ctx, lp::L_MSG + 2 * l, E, s.xm[l], s.gze, s.gbf, s.gxe));
        hipLaunchKernelGGL(lp_scatter_h_kernel, dim3(nblk(NH, 256)), dim3(256), 0, ctx->stream, N, ctx->d_lrow,
                           ctx->d_urow, ctx->d_upos, s.gxe, s.gh, s.gh2);
        std::swap(s.gh, s.gh2);
        hipLaunchKernelGGL(lp_edge_resid_kernel, dim3(nblk(EH, 256)), dim3(256), 0, ctx->stream, E, s.gxe, s.ge,
                           s.gm);
        std::swap(s.ge, s.gm);
    }
    // encoders
    hipLaunchKernelGGL(lp_act_bwd_kernel, dim3(nblk(NH, 256)), dim3(256), 0, ctx->stream, NH, s.gh, s.zn0, s.gzn,
                       s.gbf);
    LRC(layer_bwd(ctx, lp::L_ENC_N, N, s.xn0, s.gzn, s.gbf, nullptr));
    if (E)
        hipLaunchKernelGGL(lp_act_bwd_kernel, dim3(nblk(EH, 256)), dim3(256), 0, ctx->stream, EH, s.ge, s.ze0,
                           s.gze, s.gbf);
    LRC(layer_bwd(ctx, lp::L_ENC_E, E, s.xe0, s.gze, s.gbf, nullptr));
        FSD_HIP(hipGetLastError());
    return 0;
}

// Training loss of nsamp samples (x_s, b_s = A x_s, 17N each, consecutive) and,
// when grad, the gradient of every network parameter into d_grad.
// loss = sum_s |g_s - (b_F - x_F)_s|^2 + |g_s - x_U,s|^2 over sum_s |x_s|^2,
// g_s = G_k (b_F - b_U)_s.  lp_forward(train = true) must have run.
int lp_loss_grad(fsd_ctx *ctx, int nsamp, const float *d_x, const float *d_b, double *loss, bool grad)
{
    LearnedState &s = *ctx->lp;
    const int N = ctx->N, k = ctx->jac_k;
    const int64_t E = ctx->pch.nlow, n6 = 6 * (int64_t)N, n17 = 17 * (int64_t)N;
    if (nsamp <= 0 || nsamp * (2 * k + 2) > 64) { ctx->err = "learned loss: nsamp x (2 k + 2) must be <= 64"; return FSD_E_PARAM; }
    // per sample: t, v_0..v_k, u_0..u_k-1, x_0..x_k, s_0..s_k-1, and the adjoints vb_0..vb_k,
    // ub_0..ub_k-1, xb_0..xb_k, sb_0..sb_k-1
    const int per = 1 + (k + 1) + k + (k + 1) + k + (k + 1) + k + (k + 1) + k;
    const int64_t need = (int64_t)nsamp * per * n6;
    if (need > s.jv_cap) {
        FSD_HIP(hipStreamSynchronize(ctx->stream));
        LRC(lalloc(ctx, s.jv, need));
        s.jv_cap = need;
    }
    const dim3 g(row_blocks(N)), b(256);
    FSD_HIP(hipMemsetAsync(s.d_loss, 0, 2 * sizeof(double), ctx->stream));
    hipLaunchKernelGGL(lj_sumsq_kernel, dim3(256), dim3(256), 0, ctx->stream, (int64_t)nsamp * n17, d_x, s.d_loss);
    GTerms tw{}, td{};
    for (int smp = 0; smp < nsamp; ++smp) {
        float *base = s.jv + (int64_t)smp * per * n6;
        auto vec = [&](int idx) { return base + (int64_t)idx * n6; };
        float *t = vec(0);
        float *V[64], *U[64], *X[64], *S[64], *Vb[64], *Ub[64], *Xb[64], *Sb[64];
        int o = 1;
        for (int j = 0; j <= k; ++j) V[j] = vec(o++);
        for (int j = 0; j < k; ++j) U[j] = vec(o++);
        for (int j = 0; j <= k; ++j) X[j] = vec(o++);
        for (int j = 0; j < k; ++j) S[j] = vec(o++);
        for (int j = 0; j <= k; ++j) Vb[j] = vec(o++);
        for (int j = 0; j < k; ++j) Ub[j] = vec(o++);
        for (int j = 0; j <= k; ++j) Xb[j] = vec(o++);
        for (int j = 0; j < k; ++j) Sb[j] = vec(o++);
        const float *xs = d_x + smp * n17, *bs = d_b + smp * n17;
        // forward: t = P (b_F - b_U), v_0 = Dinv t, v_j+1 = Dinv (t - W v_j); x_0 = v_k,
        // x_j+1 = v_k - Dinv W^T x_j
        hipLaunchKernelGGL(lj_pre_kernel, g, b, 0, ctx->stream, N, ctx->d_perm, s.Dinv, bs, bs + 11 * (int64_t)N, t, V[0]);
        for (int j = 0; j < k; ++j)
            hipLaunchKernelGGL(lj_fwd_kernel, g, b, 0, ctx->stream, N, ctx->d_lrow, ctx->d_lcol, s.W, s.Dinv, t, V[j],
                               U[j], V[j + 1]);
        FSD_HIP(hipMemcpyAsync(X[0], V[k], sizeof(float) * n6, hipMemcpyDeviceToDevice, ctx->stream));
        for (int j = 0; j < k; ++j)
            hipLaunchKernelGGL(lj_bwd_kernel, g, b, 0, ctx->stream, N, ctx->d_urow, ctx->d_ucol, ctx->d_upos, s.W,
                               s.Dinv, V[k], X[j], S[j], X[j + 1]);
        hipLaunchKernelGGL(lj_loss_kernel, g, b, 0, ctx->stream, N, ctx->d_perm, X[k], xs, bs, s.d_loss, Xb[k]);
        if (!grad) continue;
        // reverse of the backward iterations: sb_j = -Dinv xb_j+1, xb_j = W sb_j; vb_k = sum_j xb_j
        for (int j = k - 1; j >= 0; --j) {
            hipLaunchKernelGGL(lj_dmul_kernel, g, b, 0, ctx->stream, N, s.Dinv, Xb[j + 1], -1.0f, Sb[j]);
            hipLaunchKernelGGL(lj_lower_kernel, g, b, 0, ctx->stream, N, ctx->d_lrow, ctx->d_lcol, s.W, Sb[j], 1.0f,
                               Xb[j]);
        }
        FSD_HIP(hipMemcpyAsync(Vb[k], Xb[0], sizeof(float) * n6, hipMemcpyDeviceToDevice, ctx->stream));
        for (int j = 1; j <= k; ++j)
            hipLaunchKernelGGL(lj_add_kernel, dim3(nblk(n6, 256)), dim3(256), 0, ctx->stream, n6, Xb[j], Vb[k]);
        // reverse of the forward iterations: ub_j = Dinv vb_j+1, vb_j = -W^T ub_j
        for (int j = k - 1; j >= 0; --j) {
            hipLaunchKernelGGL(lj_dmul_kernel, g, b, 0, ctx->stream, N, s.Dinv, Vb[j + 1], 1.0f, Ub[j]);
            hipLaunchKernelGGL(lj_upper_kernel, g, b, 0, ctx->stream, N, ctx->d_urow, ctx->d_ucol, ctx->d_upos, s.W,
                               Ub[j], -1.0f, Vb[j]);
        }
        // terms of gW (e = (P, Q)): -ub_j[P] v_j[Q]^T + x_j[P] sb_j[Q]^T
        for (int j = 0; j < k; ++j) {
            tw.A[tw.n] = Ub[j]; tw.B[tw.n] = V[j]; tw.coef[tw.n++] = -1.0f;
            tw.A[tw.n] = X[j]; tw.B[tw.n] = Sb[j]; tw.coef[tw.n++] = 1.0f;
        }
        // terms of gDinv: vb_0 t^T + sum_j vb_j+1 u_j^T - xb_j+1 s_j^T
        td.A[td.n] = Vb[0]; td.B[td.n] = t; td.coef[td.n++] = 1.0f;
        for (int j = 0; j < k; ++j) {
            td.A[td.n] = Vb[j + 1]; td.B[td.n] = U[j]; td.coef[td.n++] = 1.0f;
            td.A[td.n] = Xb[j + 1]; td.B[td.n] = S[j]; td.coef[td.n++] = -1.0f;
        }
    }
    if (grad) {
        if (E) hipLaunchKernelGGL(lj_gw_kernel, dim3(nblk(E * 6, 256)), dim3(256), 0, ctx->stream, E, s.erow, ctx->d_lcol, tw, s.gW);
        hipLaunchKernelGGL(lj_gd_kernel, dim3(nblk((int64_t)N * 6, 256)), dim3(256), 0, ctx->stream, N, td, s.gDinv);
        LRC(lp_backward(ctx, false));
    }
    FSD_HIP(hipGetLastError());
    FSD_HIP(hipMemcpyAsync(s.h_loss, s.d_loss, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FSD_HIP(hipStreamSynchronize(ctx->stream));
    *loss = s.h_loss[0];
    return 0;
}

// Supervised factor loss: the learned factor's pivots and pair blocks against IC(0)'s
// block LDL^T on the same pattern (d_Lblk, d_Ldiag must hold IC(0)'s factor), relative
// errors averaged over particles and over pairs; with grad, every parameter's gradient.
// lp_forward(train = true) must have run.
int lp_factor_loss_grad(fsd_ctx *ctx, double *loss, bool grad)
{
    LearnedState &s = *ctx->lp;
    const int N = ctx->N;
    const int64_t E = ctx->pch.nlow;
    FSD_HIP(hipMemsetAsync(s.d_loss, 0, 2 * sizeof(double), ctx->stream));
    hipLaunchKernelGGL(lp_pivot_target_kernel, dim3(nblk(N, 128)), dim3(128), 0, ctx->stream, N, ctx->d_Ldiag, s.D,
                       s.d_loss, s.gDinv);
    if (E)
        hipLaunchKernelGGL(lp_pair_target_kernel, dim3(nblk(E, 128)), dim3(128), 0, ctx->stream, E, ctx->d_lcol,
                           ctx->d_Lblk, ctx->d_Ldiag, ctx->d_Ablk, s.W, s.d_loss, s.gW);
    FSD_HIP(hipGetLastError());
    if (grad) LRC(lp_backward(ctx, true));
    FSD_HIP(hipMemcpyAsync(s.h_loss, s.d_loss, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FSD_HIP(hipStreamSynchronize(ctx->stream));
    *loss = s.h_loss[0];
    return 0;
}

// undo the last Adam step (the parameters and moments it started from)
int lp_revert(fsd_ctx *ctx)
{
    LearnedState &s = *ctx->lp;
    if (!s.d_wbak) { ctx->err = "no Adam step to revert"; return FSD_E_STATE; }
    const size_t b = sizeof(float) * s.nparam;
    FSD_HIP(hipMemcpyAsync(s.d_w, s.d_wbak, b, hipMemcpyDeviceToDevice, ctx->stream));
    FSD_HIP(hipMemcpyAsync(s.d_m, s.d_wbak + s.nparam, b, hipMemcpyDeviceToDevice, ctx->stream));
    FSD_HIP(hipMemcpyAsync(s.d_v, s.d_wbak + 2 * s.nparam, b, hipMemcpyDeviceToDevice, ctx->stream));
    LRC(pack_weights(ctx));
    FSD_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int lp_adam(fsd_ctx *ctx, float lr, int step)
{
    LearnedState &s = *ctx->lp;
    if (!s.d_wbak) LRC(lalloc(ctx, s.d_wbak, 3 * s.nparam));
    {
        const size_t b = sizeof(float) * s.nparam;
        FSD_HIP(hipMemcpyAsync(s.d_wbak, s.d_w, b, hipMemcpyDeviceToDevice, ctx->stream));
        FSD_HIP(hipMemcpyAsync(s.d_wbak + s.nparam, s.d_m, b, hipMemcpyDeviceToDevice, ctx->stream));
        FSD_HIP(hipMemcpyAsync(s.d_wbak + 2 * s.nparam, s.d_v, b, hipMemcpyDeviceToDevice, ctx->stream));
    }
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(lp_adam_kernel, dim3(nblk(s.nparam, 256)), dim3(256), 0, ctx->stream, s.nparam, s.d_w,
                       s.d_grad, s.d_m, s.d_v, lr, b1, b2, 1e-8f, c1, c2);
    LRC(pack_weights(ctx));
    FSD_HIP(hipStreamSynchronize(ctx->stream));
    s.have_w = true;
    return 0;
}

// safeguard levels of the last factor's pivots, [NPASS][N]
int lp_levels(fsd_ctx *ctx, int32_t *h_levels, int64_t n)
{
    LearnedState &s = *ctx->lp;
    if (n != (int64_t)lp::NPASS * ctx->N) {
        ctx->err = "learned levels: expected " + std::to_string((int64_t)lp::NPASS * ctx->N) + " values";
        return FSD_E_PARAM;
    }
    FSD_HIP(hipMemcpyAsync(h_levels, s.lvlp, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
    FSD_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

// network outputs and the factor's blocks of the current pattern, to host buffers (each may be null)
int lp_export(fsd_ctx *ctx, float *h_oute, float *h_outn, float *h_W, float *h_Dinv)
{
    LearnedState &s = *ctx->lp;
    const int64_t E = ctx->pch.nlow, N = ctx->N;
    if (h_oute && E) FSD_HIP(hipMemcpyAsync(h_oute, s.oute, sizeof(float) * E * lp::OE, hipMemcpyDeviceToHost, ctx->stream));
    if (h_outn) FSD_HIP(hipMemcpyAsync(h_outn, s.outn, sizeof(float) * N * lp::ON, hipMemcpyDeviceToHost, ctx->stream));
    if (h_W && E) FSD_HIP(hipMemcpyAsync(h_W, s.W, sizeof(float) * E * 36, hipMemcpyDeviceToHost, ctx->stream));
    if (h_Dinv) FSD_HIP(hipMemcpyAsync(h_Dinv, s.Dinv, sizeof(float) * N * 36, hipMemcpyDeviceToHost, ctx->stream));
    FSD_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

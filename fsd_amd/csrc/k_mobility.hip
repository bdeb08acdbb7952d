// PSE far-field grand mobility on MI355X.
//
//   wave space:  particle records -> tile-gather spreading (no global atomics)
//                -> rocFFT R2C (11 batched fields) -> Green's-function scaling on
//                the half spectrum -> rocFFT C2R -> per-particle contraction;
//   real space:  table-interpolated FTS pair sum, 16 lanes per particle, with an
//                epilogue that adds the wave-space result, converts the velocity
//                gradient to (W, E) and writes the generalised velocity.
// Restates Mobility.cu:131-1542 and Helper_Mobility.cu:88-410.
#include "fsd_internal.h"

#define FSD_TILE 8

// ------------------------------------------------------------- k vectors
// Mobility_SetGridk_kernel (Helper_Mobility.cu:351-410) on the R2C half
// spectrum: index (i, j, k) with k in [0, Nz/2].
__global__ void gridk_kernel(float4 *gk, int Nx, int Ny, int Nz, BoxF b, float xi, float eta, float twopi_k)
{
    const int Nzh = Nz / 2 + 1;
    const int64_t Mh = (int64_t)Nx * Ny * Nzh;
    int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= Mh) return;
    int i = (int)(tid / ((int64_t)Ny * Nzh));
    int j = (int)((tid / Nzh) % Ny);
    int k = (int)(tid % Nzh);
    gk[tid] = gridk_value(i, j, k, Nx, Ny, Nz, b, xi, eta, twopi_k);
}

// 2 pi of the k vectors: the reference's constant 3.1416926536 with parity
// (Helper_Mobility.cu:389-391)
float k_twopi(const fsd_ctx *ctx)
{
    return (float)(2.0 * (ctx->p.parity ? 3.1416926536 : 3.14159265358979323846));
}

int k_gridk(fsd_ctx *ctx)
{
    hipLaunchKernelGGL(gridk_kernel, dim3(nblk(ctx->Mh, 256)), dim3(256), 0, ctx->stream, ctx->d_gridk, ctx->pse.Nx,
                       ctx->pse.Ny, ctx->pse.Nz, ctx->box, ctx->p.xi, ctx->pse.eta, k_twopi(ctx));
    FSD_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------- spreading set-up (per step)
// Stencil origin exactly as Mobility.cu:203-225 (float fraction, truncation);
// particles are then counting-sorted by the grid cell of their (wrapped)
// origin, ascending particle id inside a cell, so spreading and contraction
// can find the particles of any block of origin cells as contiguous runs.
__global__ void stencil_origin_kernel(const float4 *pos, int N, BoxF b, int Nx, int Ny, int Nz, int P, int *orig,
                                      int *ocell, int *ocount)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float4 p = pos[i];
    float fx, fy, fz;
    box_make_fraction(b, p.x, p.y, p.z, fx, fy, fz);
    fx *= (float)Nx; fy *= (float)Ny; fz *= (float)Nz;
    int x = (int)fx, y = (int)fy, z = (int)fz;
    const int Pd2 = P / 2;
    int ox = x - Pd2 + 1 - (P % 2) * ((fx - (float)x) < 0.5f);
    int oy = y - Pd2 + 1 - (P % 2) * ((fy - (float)y) < 0.5f);
    int oz = z - Pd2 + 1 - (P % 2) * ((fz - (float)z) < 0.5f);
    orig[3 * i] = ox; orig[3 * i + 1] = oy; orig[3 * i + 2] = oz;
    int wx = ((ox % Nx) + Nx) % Nx, wy = ((oy % Ny) + Ny) % Ny, wz = ((oz % Nz) + Nz) % Nz;
    const int c = (wx * Ny + wy) * Nz + wz;
    ocell[i] = c;
    atomicAdd(&ocount[c], 1);
}

__global__ void ocell_scatter_kernel(const int *ocell, int N, const int *start, int *fill, int *list)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int c = ocell[i];
    list[start[c] + atomicAdd(&fill[c], 1)] = i;
}

__global__ void ocell_sort_kernel(const int *start, int ncell, int *list)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int s = start[c], e = start[c + 1];
    for (int a = s + 1; a < e; ++a) {
        int key = list[a], q = a - 1;
        while (q >= s && list[q] > key) { list[q + 1] = list[q]; --q; }
        list[q + 1] = key;
    }
}

int scan_excl_large(fsd_ctx *ctx, const int *in, int *out, int n, int *tmp);

int k_setup_spread(fsd_ctx *ctx, const float4 *pos)
{
    const int N = ctx->N, M = ctx->M;
    FSD_HIP(hipMemsetAsync(ctx->d_ocount, 0, sizeof(int) * M, ctx->stream));
    hipLaunchKernelGGL(stencil_origin_kernel, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, pos, N, ctx->box,
                       ctx->pse.Nx, ctx->pse.Ny, ctx->pse.Nz, ctx->pse.P, ctx->d_porig, ctx->d_ocell, ctx->d_ocount);
    if (scan_excl_large(ctx, ctx->d_ocount, ctx->d_ostart, M, ctx->d_scan_tmp)) return FSD_E_HIP;
    FSD_HIP(hipMemsetAsync(ctx->d_ocount, 0, sizeof(int) * M, ctx->stream));
    hipLaunchKernelGGL(ocell_scatter_kernel, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, ctx->d_ocell, N,
                       ctx->d_ostart, ctx->d_ocount, ctx->d_olist);
    hipLaunchKernelGGL(ocell_sort_kernel, dim3(nblk(M, 256)), dim3(256), 0, ctx->stream, ctx->d_ostart, M,
                       ctx->d_olist);
    FSD_HIP(hipGetLastError());
    ctx->spread_valid = true;
    ctx->spread_pos = (const float *)pos;
    return 0;
}

// ------------------------------------------------- force + couplet records
// per particle: F (3) and the 8 stored couplet components (Mobility_TS2C_kernel,
// Helper_Mobility.cu:294-336) packed as 3 float4
__global__ void fc_records_kernel(const float *gF, int N, float4 *rec)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float *f = gF + 6 * (int64_t)i;
    const float *s = gF + 6 * (int64_t)N + 5 * (int64_t)i;
    float Lx = -f[3], Ly = -f[4], Lz = -f[5];
    float Sxx = s[0], Sxy = s[1], Sxz = s[2], Syz = s[3], Syy = s[4];
    float C0 = Sxx, C1 = Sxy + 0.5f * Lz, C2 = Sxz - 0.5f * Ly, C3 = Syz + 0.5f * Lx;
    float C4 = Syy, C5 = Sxy - 0.5f * Lz, C6 = Sxz + 0.5f * Ly, C7 = Syz - 0.5f * Lx;
    rec[3 * (int64_t)i] = make_float4(f[0], f[1], f[2], C0);
    rec[3 * (int64_t)i + 1] = make_float4(C1, C2, C3, C4);
    rec[3 * (int64_t)i + 2] = make_float4(C5, C6, C7, 0.f);
}

// ------------------------------------------------------------ spreading
struct SpreadGeom {
    int Nx, Ny, Nz, P, tile, tnx, tny, tnz;
    float hx, hy, hz, Lx2, Ly2, Lz2, xy;
    float prefac, expfac;
};

// The origin cells whose stencils reach a tile [t0, t0 + tile) on one axis are
// t0 - P + 1 .. t0 + tile - 1 (wrapped): `halo_axis` lists them once each
// (fewer when the grid is shorter than the list) and returns the count.
__device__ int halo_axis(int t0, int tile, int P, int n, int *list)
{
    int cnt = 0;
    for (int o = t0 - P + 1; o <= t0 + tile - 1; ++o) {
        const int w = ((o % n) + n) % n;
        bool seen = false;
        for (int q = 0; q < cnt; ++q) seen |= (list[q] == w);
        if (!seen) list[cnt++] = w;
    }
    return cnt;
}

// One workgroup per tile of TILE^3 grid points, one thread per point.  The
// particles whose origin lies in the tile's halo of origin cells (the exact
// set whose stencils can reach the tile) are found as runs of the
// origin-cell-sorted list and streamed through LDS in chunks; every point
// accumulates its 11 fields in registers and is written once (no atomics).
#define SP_CH 256
#define SP_MAXAX 32
__global__ void __launch_bounds__(512) spread_kernel(SpreadGeom g, const float4 *pos, const int *orig,
                                                     const float4 *rec, const int *ostart, const int *olist,
                                                     float *grid, int64_t M)
{
    __shared__ int s_ax[3][SP_MAXAX];
    __shared__ int s_nax[3];
    __shared__ int s_zrun[SP_MAXAX][2];   // z runs: [first cell, one past last cell)
    __shared__ int s_nzrun;
    __shared__ int s_seg_beg[SP_MAXAX * SP_MAXAX * 2];
    __shared__ int s_seg_off[SP_MAXAX * SP_MAXAX * 2 + 1];
    __shared__ int s_nseg;
    __shared__ int s_o[SP_CH][3];
    __shared__ float s_p[SP_CH][3];
    __shared__ float4 s_r[SP_CH][3];

    const int tile_id = blockIdx.x;
    const int tz_i = tile_id % g.tnz, ty_i = (tile_id / g.tnz) % g.tny, tx_i = tile_id / (g.tnz * g.tny);
    const int x0 = tx_i * g.tile, y0 = ty_i * g.tile, z0 = tz_i * g.tile;
    const int lt = threadIdx.x;
    const int lx = lt / (g.tile * g.tile), ly = (lt / g.tile) % g.tile, lz = lt % g.tile;
    const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
    const bool inside = (gx < g.Nx) && (gy < g.Ny) && (gz < g.Nz);

    if (lt < 3) {
        const int t0 = (lt == 0) ? x0 : (lt == 1) ? y0 : z0;
        const int n = (lt == 0) ? g.Nx : (lt == 1) ? g.Ny : g.Nz;
        s_nax[lt] = halo_axis(t0, g.tile, g.P, n, s_ax[lt]);
    }
    __syncthreads();
    if (lt == 0) {
        // group the z cells into runs of consecutive indices
        int nr = 0;
        for (int q = 0; q < s_nax[2]; ++q) {
            const int z = s_ax[2][q];
            if (nr > 0 && s_zrun[nr - 1][1] == z) s_zrun[nr - 1][1] = z + 1;
            else { s_zrun[nr][0] = z; s_zrun[nr][1] = z + 1; ++nr; }
        }
        s_nzrun = nr;
        s_nseg = s_nax[0] * s_nax[1] * nr;
    }
    __syncthreads();
    const int nseg = s_nseg, nzr = s_nzrun, ny = s_nax[1];
    for (int sg = lt; sg < nseg; sg += blockDim.x) {
        const int zr = sg % nzr, xy = sg / nzr;
        const int x = s_ax[0][xy / ny], y = s_ax[1][xy % ny];
        const int cbase = (x * g.Ny + y) * g.Nz;
        const int b0 = ostart[cbase + s_zrun[zr][0]], b1 = ostart[cbase + s_zrun[zr][1]];
        s_seg_beg[sg] = b0;
        s_seg_off[sg + 1] = b1 - b0;
    }
    __syncthreads();
    if (lt == 0) {
        s_seg_off[0] = 0;
        for (int sg = 0; sg < nseg; ++sg) s_seg_off[sg + 1] += s_seg_off[sg];
    }
    __syncthreads();
    const int total = s_seg_off[nseg];

    float acc[11];
#pragma unroll
    for (int c = 0; c < 11; ++c) acc[c] = 0.f;

    for (int base = 0; base < total; base += SP_CH) {
        const int nch = min(SP_CH, total - base);
        __syncthreads();
        if (lt < nch) {
            const int q = base + lt;
            int lo = 0, hi = nseg - 1;          // segment sg with off[sg] <= q < off[sg + 1]
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_seg_off[mid] <= q) lo = mid; else hi = mid - 1;
            }
            const int i = olist[s_seg_beg[lo] + (q - s_seg_off[lo])];
            s_o[lt][0] = orig[3 * i]; s_o[lt][1] = orig[3 * i + 1]; s_o[lt][2] = orig[3 * i + 2];
            const float4 p = pos[i];
            s_p[lt][0] = p.x; s_p[lt][1] = p.y; s_p[lt][2] = p.z;
            s_r[lt][0] = rec[3 * (int64_t)i]; s_r[lt][1] = rec[3 * (int64_t)i + 1]; s_r[lt][2] = rec[3 * (int64_t)i + 2];
        }
        __syncthreads();
        if (inside) {
            for (int q = 0; q < nch; ++q) {
                int tx = gx - s_o[q][0]; tx += (tx < 0) ? g.Nx : 0; tx -= (tx >= g.Nx) ? g.Nx : 0;
                int ty = gy - s_o[q][1]; ty += (ty < 0) ? g.Ny : 0; ty -= (ty >= g.Ny) ? g.Ny : 0;
                int tz = gz - s_o[q][2]; tz += (tz < 0) ? g.Nz : 0; tz -= (tz >= g.Nz) ? g.Nz : 0;
                if (tx < g.P && ty < g.P && tz < g.P) {
                    // unwrapped grid node of this stencil point, relative to the particle
                    const float uy = g.hy * (float)(s_o[q][1] + ty) - g.Ly2;
                    const float dx = g.hx * (float)(s_o[q][0] + tx) - g.Lx2 + g.xy * uy - s_p[q][0];
                    const float dy = uy - s_p[q][1];
                    const float dz = g.hz * (float)(s_o[q][2] + tz) - g.Lz2 - s_p[q][2];
                    const float w = g.prefac * __expf(-g.expfac * (dx * dx + dy * dy + dz * dz));
                    const float4 r0 = s_r[q][0], r1 = s_r[q][1], r2 = s_r[q][2];
                    acc[0] += w * r0.x; acc[1] += w * r0.y; acc[2] += w * r0.z;
                    acc[3] += w * r0.w; acc[4] += w * r1.x; acc[5] += w * r1.y; acc[6] += w * r1.z;
                    acc[7] += w * r1.w; acc[8] += w * r2.x; acc[9] += w * r2.y; acc[10] += w * r2.z;
                }
            }
        }
    }
    if (inside) {
        const int64_t id = ((int64_t)gx * g.Ny + gy) * g.Nz + gz;
#pragma unroll
        for (int c = 0; c < 11; ++c) grid[c * M + id] = acc[c];
    }
}

// Spreading for the usual stencil (P = 6, tile 8) with the work of a grid
// point limited to its own stencil window.  The tile's H^3 halo of origin
// cells (H = tile + P - 1) is counted and scanned in LDS, its particles are
// staged in LDS in (x, y, z) halo-cell order, and a point at local (lx, ly,
// lz) visits, for each of its 36 halo columns (lx + a, ly + b), the
// consecutive staged particles of cells lz .. lz + P - 1: exactly the
// particles whose stencil covers it.  A tile whose halo holds more than
// SP2_CAP particles reads them from global memory instead (about 190 are staged
// at volume fraction 0.45; the cap keeps the LDS at 50 KB, three workgroups per
// CU).  Same weights and accumulation as spread_kernel (each point sums its
// particles in the same cell order, and writes once).
#define SP2_CAP 512
// contribution of one particle (origin o, position p, records r) to grid point (gx, gy, gz)
__device__ __forceinline__ void spread_point_add(const SpreadGeom &g, int gx, int gy, int gz, int o0, int o1, int o2,
                                                 float p0, float p1, float p2, float4 r0, float4 r1, float4 r2,
                                                 float *acc)
{
    int tx = gx - o0; tx += (tx < 0) ? g.Nx : 0; tx -= (tx >= g.Nx) ? g.Nx : 0;
    int ty = gy - o1; ty += (ty < 0) ? g.Ny : 0; ty -= (ty >= g.Ny) ? g.Ny : 0;
    int tz = gz - o2; tz += (tz < 0) ? g.Nz : 0; tz -= (tz >= g.Nz) ? g.Nz : 0;
    const float uy = g.hy * (float)(o1 + ty) - g.Ly2;
    const float dx = g.hx * (float)(o0 + tx) - g.Lx2 + g.xy * uy - p0;
    const float dy = uy - p1;
    const float dz = g.hz * (float)(o2 + tz) - g.Lz2 - p2;
    const float w = g.prefac * __expf(-g.expfac * (dx * dx + dy * dy + dz * dz));
    acc[0] += w * r0.x; acc[1] += w * r0.y; acc[2] += w * r0.z;
    acc[3] += w * r0.w; acc[4] += w * r1.x; acc[5] += w * r1.y; acc[6] += w * r1.z;
    acc[7] += w * r1.w; acc[8] += w * r2.x; acc[9] += w * r2.y; acc[10] += w * r2.z;
}

template <int TILE, int PP>
__global__ void __launch_bounds__(512) spread_tile2_kernel(SpreadGeom g, const float4 *pos, const int *orig,
                                                           const float4 *rec, const int *ostart, const int *olist,
                                                           float *grid, int64_t M)
{
    constexpr int H = TILE + PP - 1;
    constexpr int NQ = H * H * H;                // halo cells
    constexpr int PER = (NQ + 511) / 512;
    __shared__ int s_beg[NQ];                    // first olist entry of each halo cell
    __shared__ int s_base[NQ + 1];               // exclusive scan of the cells' counts
    __shared__ int s_wsum[8];
    // staged particle: the offsets (cx, cy, cy - py, cz) of its stencil in the tile's frame and its
    // records scaled by the Gaussian's prefactor
    __shared__ float4 s_c[SP2_CAP];
    __shared__ float4 s_r[SP2_CAP][3];

    const int tile_id = xcd_remap(blockIdx.x, gridDim.x);
    const int tz_i = tile_id % g.tnz, ty_i = (tile_id / g.tnz) % g.tny, tx_i = tile_id / (g.tnz * g.tny);
    const int x0 = tx_i * TILE, y0 = ty_i * TILE, z0 = tz_i * TILE;
    const int lt = threadIdx.x;
    // counts of the halo cells, PER consecutive cells per thread
    int cnt[PER];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int q = lt * PER + k;
        cnt[k] = 0;
        if (q < NQ) {
            const int hx = q / (H * H), hy = (q / H) % H, hz = q % H;
            int x = x0 - (PP - 1) + hx, y = y0 - (PP - 1) + hy, z = z0 - (PP - 1) + hz;
            x += (x < 0) ? g.Nx : 0; x -= (x >= g.Nx) ? g.Nx : 0; x -= (x >= g.Nx) ? g.Nx : 0;
            y += (y < 0) ? g.Ny : 0; y -= (y >= g.Ny) ? g.Ny : 0; y -= (y >= g.Ny) ? g.Ny : 0;
            z += (z < 0) ? g.Nz : 0; z -= (z >= g.Nz) ? g.Nz : 0; z -= (z >= g.Nz) ? g.Nz : 0;
            const int c = (x * g.Ny + y) * g.Nz + z;
            const int b = ostart[c];
            s_beg[q] = b;
            cnt[k] = ostart[c + 1] - b;
        }
        sum += cnt[k];
    }
    // exclusive scan over the threads: wave inclusive scan, then wave offsets
    const int lane = lt & 63, wv = lt >> 6;
    int inc = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
    }
    if (lane == 63) s_wsum[wv] = inc;
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < wv; ++w) woff += s_wsum[w];
    int run = woff + inc - sum;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int q = lt * PER + k;
        if (q < NQ) s_base[q] = run;
        run += cnt[k];
    }
    int total = 0;
    for (int w = 0; w < 8; ++w) total += s_wsum[w];
    if (lt == 0) s_base[NQ] = total;
    const bool staged = total <= SP2_CAP;
    __syncthreads();
    if (staged) {
        // staged slots sl = lt and lt + 512 (at most SP2_CAP of them): the halo cell holding the
        // slot by binary search over the scan (the last cell whose base is <= sl is non-empty and
        // holds it), then the particle's index and records; the loads of both slots overlap
        static_assert(SP2_CAP <= 2 * 512, "two staged slots per thread");
        int q2[2], sl2[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int sl = lt + h * 512;
            int lo = 0, hi = NQ - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_base[mid] <= sl) lo = mid; else hi = mid - 1;
            }
            q2[h] = lo;
            sl2[h] = sl;
        }
        int i2[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const bool ok = sl2[h] < total;
            i2[h] = olist[ok ? s_beg[q2[h]] + (sl2[h] - s_base[q2[h]]) : 0];
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int sl = sl2[h], i = i2[h];
            if (sl < total) {
                // the particle's origin is the halo cell (hc) up to the period, so a point at local
                // (lx, ly, lz) of its stencil is node o + (P - 1) - hc + l: with the node's position
                // h (o + P - 1 - hc) - L/2 + h l, its offset from the particle is c + h l
                const int q = q2[h];
                const int hxc = q / (H * H), hyc = (q / H) % H, hzc = q % H;
                const int ox = orig[3 * i], oy = orig[3 * i + 1], oz = orig[3 * i + 2];
                const float4 p = pos[i];
                const float cx = g.hx * (float)(ox + (PP - 1) - hxc) - g.Lx2 - p.x;
                const float cy = g.hy * (float)(oy + (PP - 1) - hyc) - g.Ly2;
                const float cz = g.hz * (float)(oz + (PP - 1) - hzc) - g.Lz2 - p.z;
                s_c[sl] = make_float4(cx, cy, cy - p.y, cz);
                const float4 r0 = rec[3 * (int64_t)i], r1 = rec[3 * (int64_t)i + 1], r2 = rec[3 * (int64_t)i + 2];
                const float f = g.prefac;
                s_r[sl][0] = make_float4(f * r0.x, f * r0.y, f * r0.z, f * r0.w);
                s_r[sl][1] = make_float4(f * r1.x, f * r1.y, f * r1.z, f * r1.w);
                s_r[sl][2] = make_float4(f * r2.x, f * r2.y, f * r2.z, 0.f);
            }
        }
    }
    __syncthreads();
    const int lx = lt / (TILE * TILE), ly = (lt / TILE) % TILE, lz = lt % TILE;
    const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
    if (lt >= TILE * TILE * TILE || gx >= g.Nx || gy >= g.Ny || gz >= g.Nz) return;
    float acc[11];
#pragma unroll
    for (int c = 0; c < 11; ++c) acc[c] = 0.f;
    // one loop per source (separate code, so the staged loop reads LDS only)
    if (staged) {
        // offsets of this point in the tile's frame; a staged particle's are added to them
        const float ax = g.hx * (float)lx, ay = g.hy * (float)ly, az = g.hz * (float)lz;
        const float nexp = -g.expfac;
        for (int a = 0; a < PP; ++a)
            for (int b = 0; b < PP; ++b) {
                const int q0 = ((lx + a) * H + (ly + b)) * H + lz;
                const int s1 = s_base[q0 + PP];
                for (int sl = s_base[q0]; sl < s1; ++sl) {
                    const float4 c = s_c[sl];
                    const float uy = c.y + ay;
                    const float dx = c.x + ax + g.xy * uy, dy = c.z + ay, dz = c.w + az;
                    const float e = __expf(nexp * (dx * dx + dy * dy + dz * dz));
                    const float4 r0 = s_r[sl][0], r1 = s_r[sl][1], r2 = s_r[sl][2];
                    acc[0] += e * r0.x; acc[1] += e * r0.y; acc[2] += e * r0.z;
                    acc[3] += e * r0.w; acc[4] += e * r1.x; acc[5] += e * r1.y; acc[6] += e * r1.z;
                    acc[7] += e * r1.w; acc[8] += e * r2.x; acc[9] += e * r2.y; acc[10] += e * r2.z;
                }
            }
    } else {
        for (int a = 0; a < PP; ++a)
            for (int b = 0; b < PP; ++b) {
                const int q0 = ((lx + a) * H + (ly + b)) * H + lz;
                for (int q = q0; q < q0 + PP; ++q) {
                    const int beg = s_beg[q], n = s_base[q + 1] - s_base[q];
                    for (int m = 0; m < n; ++m) {
                        const int i = olist[beg + m];
                        const float4 p = pos[i];
                        spread_point_add(g, gx, gy, gz, orig[3 * i], orig[3 * i + 1], orig[3 * i + 2], p.x, p.y, p.z,
                                         rec[3 * (int64_t)i], rec[3 * (int64_t)i + 1], rec[3 * (int64_t)i + 2], acc);
                    }
                }
            }
    }
    const int64_t id = ((int64_t)gx * g.Ny + gy) * g.Nz + gz;
#pragma unroll
    for (int c = 0; c < 11; ++c) grid[c * M + id] = acc[c];
}

// ------------------------------------------------------------ Green kernel
// Mobility_WaveSpace_Green_kernel (Mobility.cu:315-479) on the half spectrum.
// Field order: Fx Fy Fz | C_xx C_xy C_xz C_yz C_yy C_yx C_zx C_zy.
__device__ inline void green_apply(const float4 tk, bool origin, const float2 v[11], float2 o[11])
{
#pragma unroll
    for (int c = 0; c < 11; ++c) o[c] = make_float2(0.f, 0.f);
    if (origin) return;
    const float k[3] = {tk.x, tk.y, tk.z};
    const float ksq = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    const float kk = sqrtf(ksq);
    // couplet matrix rows (x, y, z)
    const float2 Cm[3][3] = {{v[3], v[4], v[5]}, {v[8], v[7], v[6]}, {v[9], v[10], make_float2(-(v[3].x + v[7].x), -(v[3].y + v[7].y))}};
    float2 kdF;
    kdF.x = (k[0] * v[0].x + k[1] * v[1].x + k[2] * v[2].x) / ksq;
    kdF.y = (k[0] * v[0].y + k[1] * v[1].y + k[2] * v[2].y) / ksq;
    float2 Cdk[3];
    for (int a = 0; a < 3; ++a) {
        Cdk[a].x = Cm[a][0].x * k[0] + Cm[a][1].x * k[1] + Cm[a][2].x * k[2];
        Cdk[a].y = Cm[a][0].y * k[0] + Cm[a][1].y * k[1] + Cm[a][2].y * k[2];
    }
    float2 kCk;
    kCk.x = (Cdk[0].x * k[0] + Cdk[1].x * k[1] + Cdk[2].x * k[2]) / ksq;
    kCk.y = (Cdk[0].y * k[0] + Cdk[1].y * k[1] + Cdk[2].y * k[2]) / ksq;
    const float sk = sinf(kk), ck = cosf(kk);
    const float j0 = sk / kk;
    const float jt = 3.0f * (sk - kk * ck) / (ksq * kk);
    const float Buf = tk.w * j0 * j0;
    const float Buc = -tk.w * j0 * jt;        // UC (imaginary coupling)
    const float Bdf = tk.w * j0 * jt;         // DF (imaginary coupling)
    const float q = (sk - kk * ck) / (ksq * kk);
    const float Bdc = tk.w * 9.0f * q * q;    // DC
    for (int a = 0; a < 3; ++a) {
        float2 ff = v[a];
        float2 u;
        u.x = (ff.x - k[a] * kdF.x) * Buf;
        u.y = (ff.y - k[a] * kdF.y) * Buf;
        // + i (Cdk_a - k_a kCk) Buc
        u.x += -(Cdk[a].y - k[a] * kCk.y) * Buc;
        u.y += (Cdk[a].x - k[a] * kCk.x) * Buc;
        o[a] = u;
    }
    // gradient slots in field order: (x,x) (x,y) (x,z) (y,z) (y,y) (y,x) (z,x) (z,y)
    const int sa[8] = {0, 0, 0, 1, 1, 1, 2, 2}, sc[8] = {0, 1, 2, 2, 1, 0, 0, 1};
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int a = sa[s], c = sc[s];
        const float kk_ac = k[a] * k[c];
        float2 z;
        z.x = v[a].x * k[c] - kk_ac * kdF.x;
        z.y = v[a].y * k[c] - kk_ac * kdF.y;
        float2 d;
        d.x = -z.y * Bdf + (Cdk[a].x * k[c] - kk_ac * kCk.x) * Bdc;
        d.y = z.x * Bdf + (Cdk[a].y * k[c] - kk_ac * kCk.y) * Bdc;
        o[3 + s] = d;
    }
}

struct GreenGeom {
    int Nx, Ny, Nz;
    BoxF b;
    float xi, eta, twopi_k;
};

// the spread grid is real, so its spectrum at the conjugate mode is conj(v);
// modes with a Nyquist index get the Hermitian part of the two outputs
__global__ void green_kernel(float2 *spec, const float4 *gk, int64_t Mh, GreenGeom g)
{
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= Mh) return;
    const int Nzh = g.Nz / 2 + 1;
    const int k = (int)(id % Nzh), j = (int)((id / Nzh) % g.Ny), i = (int)(id / ((int64_t)Nzh * g.Ny));
    float2 v[11];
#pragma unroll
    for (int c = 0; c < 11; ++c) v[c] = spec[c * Mh + id];
    float2 o[11];
    green_apply(gk[id], id == 0, v, o);
    if (has_nyquist(i, j, k, g.Nx, g.Ny, g.Nz)) {
        const int ic = (g.Nx - i) % g.Nx, jc = (g.Ny - j) % g.Ny, kc = (g.Nz - k) % g.Nz;
        const float4 tkc = gridk_value(ic, jc, kc, g.Nx, g.Ny, g.Nz, g.b, g.xi, g.eta, g.twopi_k);
        float2 vc[11], oc[11];
#pragma unroll
        for (int c = 0; c < 11; ++c) vc[c] = make_float2(v[c].x, -v[c].y);
        green_apply(tkc, false, vc, oc);
#pragma unroll
        for (int c = 0; c < 11; ++c) o[c] = make_float2(0.5f * (o[c].x + oc[c].x), 0.5f * (o[c].y - oc[c].y));
    }
#pragma unroll
    for (int c = 0; c < 11; ++c) spec[c * Mh + id] = o[c];
}

// The x direction of the 3-D transform with the Green's function between its
// forward and inverse halves, on the half spectrum after the (y, z) transforms:
// element (c, x, ky, kz) at c Mh + (x Ny + ky) Nzh + kz.  One workgroup takes one
// ky and FX_KC consecutive kz for all 11 fields, so every mode it holds after the
// forward transform has all 11 fields in LDS and green_kernel's update applies
// point by point; the inverse brings the columns back to x.  Nx = 8 B: forward,
// an 8-point DFT over n1 for each n2 (x = B n1 + n2), the twiddle W^(n2 k1), then a
// B-point DFT over n2 for each k1 (fx_dft8, fx_dftB), leaving mode k1 + 8 k2 at
// B k1 + k2; the inverse
// runs the same stages backwards with conjugate twiddles (unnormalised, as the
// library's transforms).  W = exp(-2 pi i / Nx) from an LDS table built in FP64.
// small DFTs of the x pass (INV: the conjugate kernel, unnormalised): 8 points by radix 2,
// 9 points as 3 x 3, other lengths directly from the table wB[j] = W_B^j
__device__ __forceinline__ float2 fx_mul_i(float2 v, bool inv)   // v * (inv ? i : -i)
{
    return inv ? make_float2(-v.y, v.x) : make_float2(v.y, -v.x);
}
template <bool INV>
__device__ __forceinline__ void fx_dft4(const float2 c[4], float2 y[4])
{
    const float2 t0 = make_float2(c[0].x + c[2].x, c[0].y + c[2].y), t1 = make_float2(c[0].x - c[2].x, c[0].y - c[2].y);
    const float2 t2 = make_float2(c[1].x + c[3].x, c[1].y + c[3].y);
    const float2 t3 = fx_mul_i(make_float2(c[1].x - c[3].x, c[1].y - c[3].y), INV);
    y[0] = make_float2(t0.x + t2.x, t0.y + t2.y); y[2] = make_float2(t0.x - t2.x, t0.y - t2.y);
    y[1] = make_float2(t1.x + t3.x, t1.y + t3.y); y[3] = make_float2(t1.x - t3.x, t1.y - t3.y);
}
template <bool INV>
__device__ __forceinline__ void fx_dft8(const float2 x[8], float2 y[8])
{
    constexpr float R = 0.70710678118654752f;
    float2 a[4], b[4], e[4], o[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        a[n] = make_float2(x[n].x + x[n + 4].x, x[n].y + x[n + 4].y);
        b[n] = make_float2(x[n].x - x[n + 4].x, x[n].y - x[n + 4].y);
    }
    // b_n W8^(-+n): W8 = R (1 -+ i), W8^2 = -+i, W8^3 = R (-1 -+ i)
    b[1] = INV ? make_float2(R * (b[1].x - b[1].y), R * (b[1].x + b[1].y))
               : make_float2(R * (b[1].x + b[1].y), R * (b[1].y - b[1].x));
    b[2] = fx_mul_i(b[2], INV);
    b[3] = INV ? make_float2(-R * (b[3].x + b[3].y), R * (b[3].x - b[3].y))
               : make_float2(R * (b[3].y - b[3].x), -R * (b[3].x + b[3].y));
    fx_dft4<INV>(a, e);
    fx_dft4<INV>(b, o);
#pragma unroll
    for (int k = 0; k < 4; ++k) { y[2 * k] = e[k]; y[2 * k + 1] = o[k]; }
}
template <bool INV>
__device__ __forceinline__ void fx_dft3(float2 a, float2 b, float2 c, float2 &X0, float2 &X1, float2 &X2)
{
    constexpr float H = 0.86602540378443865f;   // sqrt(3) / 2
    const float2 sm = make_float2(b.x + c.x, b.y + c.y), d = make_float2(b.x - c.x, b.y - c.y);
    const float2 t = fx_mul_i(make_float2(H * d.x, H * d.y), INV);
    const float2 m = make_float2(a.x - 0.5f * sm.x, a.y - 0.5f * sm.y);
    X0 = make_float2(a.x + sm.x, a.y + sm.y);
    X1 = make_float2(m.x + t.x, m.y + t.y);
    X2 = make_float2(m.x - t.x, m.y - t.y);
}
template <int B, bool INV>
__device__ __forceinline__ void fx_dftB(const float2 x[B], float2 X[B], const float2 wB[B])
{
    auto mul = [](float2 a, float2 w) {   // a * w, or a * conj(w) for the inverse
        return INV ? make_float2(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y)
                   : make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
    };
    if constexpr (B == 9) {   // n = 3 m1 + m2, k = k1 + 3 k2
        float2 u[3][3];
#pragma unroll
        for (int m2 = 0; m2 < 3; ++m2) {
            fx_dft3<INV>(x[m2], x[3 + m2], x[6 + m2], u[0][m2], u[1][m2], u[2][m2]);
#pragma unroll
            for (int k1 = 1; k1 < 3; ++k1)
                if (m2 > 0) u[k1][m2] = mul(u[k1][m2], wB[m2 * k1]);
        }
#pragma unroll
        for (int k1 = 0; k1 < 3; ++k1) fx_dft3<INV>(u[k1][0], u[k1][1], u[k1][2], X[k1], X[k1 + 3], X[k1 + 6]);
    } else {
#pragma unroll
        for (int k = 0; k < B; ++k) {
            float2 acc = make_float2(0.f, 0.f);
#pragma unroll
            for (int n = 0; n < B; ++n) {
                const float2 p = mul(x[n], wB[(n * k) % B]);
                acc.x += p.x; acc.y += p.y;
            }
            X[k] = acc;
        }
    }
}

#define FX_KC 4
template <int B>
__global__ void __launch_bounds__(256) fft_x_green_kernel(float2 *spec, const float4 *gk, int64_t Mh, GreenGeom g)
{
    constexpr int NX = 8 * B;
    __shared__ float2 buf[11 * NX * FX_KC];   // [field][x][kz in the chunk]
    __shared__ float2 tw[NX];
    const int tid = threadIdx.x;
    const int Nzh = g.Nz / 2 + 1;
    const int nchunk = (Nzh + FX_KC - 1) / FX_KC;
    // the chunks of a ky row share cache lines (a chunk is FX_KC of the row's Nzh modes): a contiguous
    // range of rows per XCD keeps every line's readers and writers on one L2
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const int ky = blk / nchunk, kz0 = (blk % nchunk) * FX_KC;
    const int kc = min(FX_KC, Nzh - kz0);
    if (tid < NX) {
        double sn, cs;
        sincospi(-2.0 * (double)tid / (double)NX, &sn, &cs);
        tw[tid] = make_float2((float)cs, (float)sn);
    }
    const int nload = 11 * NX * kc;
    for (int i = tid; i < nload; i += 256) {
        const int c = i / (NX * kc), r = i - c * (NX * kc), x = r / kc, z = r - x * kc;
        buf[(c * NX + x) * FX_KC + z] = spec[c * Mh + ((int64_t)x * g.Ny + ky) * Nzh + kz0 + z];
    }
    __syncthreads();
    const int ncol = 11 * kc;
    // W_B^j = W^(8 j)
    float2 wB[B];
#pragma unroll
    for (int j = 0; j < B; ++j) wB[j] = tw[8 * j];
    auto at = [&](int col, int x) -> float2 & { return buf[((col / kc) * NX + x) * FX_KC + col % kc]; };
    auto cmul = [](float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); };
    auto cmulc = [](float2 a, float2 b) { return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); };
    // forward stage A: 8-point DFTs, then the twiddles
    for (int task = tid; task < ncol * B; task += 256) {
        const int col = task / B, n2 = task - col * B;
        float2 xs[8], y[8];
#pragma unroll
        for (int n1 = 0; n1 < 8; ++n1) xs[n1] = at(col, B * n1 + n2);
        fx_dft8<false>(xs, y);
#pragma unroll
        for (int k1 = 0; k1 < 8; ++k1) at(col, B * k1 + n2) = cmul(y[k1], tw[n2 * k1]);
    }
    __syncthreads();
    // forward stage B: B-point DFTs
    for (int task = tid; task < ncol * 8; task += 256) {
        const int col = task >> 3, k1 = task & 7;
        float2 ys[B], X[B];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) ys[n2] = at(col, B * k1 + n2);
        fx_dftB<B, false>(ys, X, wB);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) at(col, B * k1 + k2) = X[k2];
    }
    __syncthreads();
    // Green's function (green_kernel's update) at every mode of the block
    for (int pt = tid; pt < NX * kc; pt += 256) {
        const int kx = pt / kc, z = pt - kx * kc;
        const int pos = B * (kx & 7) + (kx >> 3);
        const int k = kz0 + z;
        const int64_t id = ((int64_t)kx * g.Ny + ky) * Nzh + k;
        float2 v[11], o[11];
#pragma unroll
        for (int c = 0; c < 11; ++c) v[c] = buf[(c * NX + pos) * FX_KC + z];
        green_apply(gk[id], id == 0, v, o);
        if (has_nyquist(kx, ky, k, g.Nx, g.Ny, g.Nz)) {
            const int ic = (g.Nx - kx) % g.Nx, jc = (g.Ny - ky) % g.Ny, kcj = (g.Nz - k) % g.Nz;
            const float4 tkc = gridk_value(ic, jc, kcj, g.Nx, g.Ny, g.Nz, g.b, g.xi, g.eta, g.twopi_k);
            float2 vc[11], oc[11];
#pragma unroll
            for (int c = 0; c < 11; ++c) vc[c] = make_float2(v[c].x, -v[c].y);
            green_apply(tkc, false, vc, oc);
#pragma unroll
            for (int c = 0; c < 11; ++c) o[c] = make_float2(0.5f * (o[c].x + oc[c].x), 0.5f * (o[c].y - oc[c].y));
        }
#pragma unroll
        for (int c = 0; c < 11; ++c) buf[(c * NX + pos) * FX_KC + z] = o[c];
    }
    __syncthreads();
    // inverse stage B, then the conjugate twiddles
    for (int task = tid; task < ncol * 8; task += 256) {
        const int col = task >> 3, k1 = task & 7;
        float2 Xs[B], y[B];
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) Xs[k2] = at(col, B * k1 + k2);
        fx_dftB<B, true>(Xs, y, wB);
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) at(col, B * k1 + n2) = cmulc(y[n2], tw[n2 * k1]);
    }
    __syncthreads();
    // inverse stage A: 8-point DFTs back to x
    for (int task = tid; task < ncol * B; task += 256) {
        const int col = task / B, n2 = task - col * B;
        float2 ys[8], x[8];
#pragma unroll
        for (int k1 = 0; k1 < 8; ++k1) ys[k1] = at(col, B * k1 + n2);
        fx_dft8<true>(ys, x);
#pragma unroll
        for (int n1 = 0; n1 < 8; ++n1) at(col, B * n1 + n2) = x[n1];
    }
    __syncthreads();
    for (int i = tid; i < nload; i += 256) {
        const int c = i / (NX * kc), r = i - c * (NX * kc), x = r / kc, z = r - x * kc;
        spec[c * Mh + ((int64_t)x * g.Ny + ky) * Nzh + kz0 + z] = buf[(c * NX + x) * FX_KC + z];
    }
}

// launch of fft_x_green_kernel for Nx = 8 B (B in 5..11, so the block fits in 64 KiB of LDS); false
// for other sizes
static bool launch_fft_x_green(fsd_ctx *ctx, const GreenGeom &gg)
{
    const int Nzh = ctx->pse.Nz / 2 + 1;
    const dim3 grid(ctx->pse.Ny * ((Nzh + FX_KC - 1) / FX_KC)), blk(256);
    float2 *spec = ctx->d_spec;
    const float4 *gk = ctx->d_gridk;
    const int64_t Mh = ctx->Mh;
    switch (ctx->fx_B) {
#define FX_CASE(b) case b: hipLaunchKernelGGL(fft_x_green_kernel<b>, grid, blk, 0, ctx->stream, spec, gk, Mh, gg); return true;
        FX_CASE(5) FX_CASE(6) FX_CASE(7) FX_CASE(8) FX_CASE(9) FX_CASE(10) FX_CASE(11)
#undef FX_CASE
        default: return false;
    }
}

// ------------------------------------------------------------ contraction
// Mobility_WaveSpace_ContractU/D (Mobility.cu:503-841): one wavefront per
// particle, lanes over the (y, z) stencil pairs, loop over x.  Output per
// particle: u (3) then the gradient in the reference's transposed slots.
__global__ void __launch_bounds__(256) contract_kernel(SpreadGeom g, const float4 *pos, const int *orig,
                                                       const float *grid, int64_t M, int N, float quadw_prefac,
                                                       float *wave)
{
    const int wave_id = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave_id >= N) return;
    const int i = wave_id;
    const int ox = orig[3 * i], oy = orig[3 * i + 1], oz = orig[3 * i + 2];
    const float4 p = pos[i];
    float acc[11];
#pragma unroll
    for (int c = 0; c < 11; ++c) acc[c] = 0.f;
    const int P = g.P;
    for (int yz = lane; yz < P * P; yz += 64) {
        const int ty = yz / P, tz = yz % P;
        int y = oy + ty; const float uy = g.hy * (float)y - g.Ly2;
        int z = oz + tz; const float dz = g.hz * (float)z - g.Lz2 - p.z;
        const float dy = uy - p.y;
        y = ((y % g.Ny) + g.Ny) % g.Ny;
        z = ((z % g.Nz) + g.Nz) % g.Nz;
        for (int tx = 0; tx < P; ++tx) {
            int x = ox + tx;
            const float dx = g.hx * (float)x - g.Lx2 + g.xy * uy - p.x;
            x = ((x % g.Nx) + g.Nx) % g.Nx;
            const float w = quadw_prefac * __expf(-g.expfac * (dx * dx + dy * dy + dz * dz));
            const int64_t id = ((int64_t)x * g.Ny + y) * g.Nz + z;
#pragma unroll
            for (int c = 0; c < 11; ++c) acc[c] += w * grid[c * M + id];
        }
    }
#pragma unroll
    for (int c = 0; c < 11; ++c)
        for (int o = 32; o > 0; o >>= 1) acc[c] += __shfl_xor(acc[c], o, 64);
    if (lane == 0) {
        float *w = wave + 11 * (int64_t)i;
        w[0] = acc[0]; w[1] = acc[1]; w[2] = acc[2];
        // fields XX XY XZ YZ YY YX ZX ZY -> slots {XX, YX, ZX, ZY, YY, XY, XZ, YZ}
        w[3] = acc[3]; w[4] = acc[8]; w[5] = acc[9]; w[6] = acc[10];
        w[7] = acc[7]; w[8] = acc[4]; w[9] = acc[5]; w[10] = acc[6];
    }
}

// Tile form of the contraction: one workgroup per tile of TILE^3 origin
// cells stages the grid block [t0, t0 + TILE + P - 1)^3 of all 11 fields in
// LDS (coalesced loads, wrapped indices), then contracts the particles whose
// origin lies in the tile, one wavefront per particle, from LDS.  Requires
// 11 (TILE + P - 1)^3 floats of LDS (TILE 8, P 6: 96.7 KB).
#define CT_S 13
__global__ void __launch_bounds__(512) contract_tile_kernel(SpreadGeom g, const float4 *pos, const int *orig,
                                                            const int *ostart, const int *olist, const float *grid,
                                                            int64_t M, float quadw_prefac, float *wave)
{
    extern __shared__ float s_g[];           // [11][S][S][S]
    __shared__ int s_run_beg[64], s_run_off[65];
    const int S = g.tile + g.P - 1;
    const int S3 = S * S * S;
    const int tile_id = xcd_remap(blockIdx.x, gridDim.x);   // a slab of tiles per XCD: halos shared in its L2
    const int tz_i = tile_id % g.tnz, ty_i = (tile_id / g.tnz) % g.tny, tx_i = tile_id / (g.tnz * g.tny);
    const int x0 = tx_i * g.tile, y0 = ty_i * g.tile, z0 = tz_i * g.tile;
    const int lt = threadIdx.x;
    // origin cells of the tile: columns (x, y), z in [z0, z1)
    const int xe = min(x0 + g.tile, g.Nx), ye = min(y0 + g.tile, g.Ny), z1 = min(z0 + g.tile, g.Nz);
    const int ncx = xe - x0, ncy = ye - y0;
    if (lt < ncx * ncy) {
        const int x = x0 + lt / ncy, y = y0 + lt % ncy;
        const int cb = (x * g.Ny + y) * g.Nz;
        const int b0 = ostart[cb + z0], b1 = ostart[cb + z1];
        s_run_beg[lt] = b0;
        s_run_off[lt + 1] = b1 - b0;
    }
    // grid block, wrapped
    for (int idx = lt; idx < 11 * S3; idx += blockDim.x) {
        const int c = idx / S3, r = idx - c * S3;
        const int ax = r / (S * S), ay = (r / S) % S, az = r % S;
        int x = x0 + ax, y = y0 + ay, z = z0 + az;
        x -= (x >= g.Nx) ? g.Nx : 0; x -= (x >= g.Nx) ? g.Nx : 0;
        y -= (y >= g.Ny) ? g.Ny : 0; y -= (y >= g.Ny) ? g.Ny : 0;
        z -= (z >= g.Nz) ? g.Nz : 0; z -= (z >= g.Nz) ? g.Nz : 0;
        s_g[idx] = grid[c * M + ((int64_t)x * g.Ny + y) * g.Nz + z];
    }
    __syncthreads();
    if (lt == 0) {
        s_run_off[0] = 0;
        for (int q = 0; q < ncx * ncy; ++q) s_run_off[q + 1] += s_run_off[q];
    }
    __syncthreads();
    const int nruns = ncx * ncy;
    const int total = s_run_off[nruns];
    const int wv = lt >> 6, lane = lt & 63, nw = blockDim.x >> 6;
    const int P = g.P;
    for (int q = wv; q < total; q += nw) {
        int lo = 0, hi = nruns - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_run_off[mid] <= q) lo = mid; else hi = mid - 1;
        }
        const int i = olist[s_run_beg[lo] + (q - s_run_off[lo])];
        const int ox = orig[3 * i], oy = orig[3 * i + 1], oz = orig[3 * i + 2];
        const float4 p = pos[i];
        // local block coordinates of the stencil's first node (wrapped origin - tile start)
        const int bx = (((ox % g.Nx) + g.Nx) % g.Nx) - x0;
        const int by = (((oy % g.Ny) + g.Ny) % g.Ny) - y0;
        const int bz = (((oz % g.Nz) + g.Nz) % g.Nz) - z0;
        float acc[11];
#pragma unroll
        for (int c = 0; c < 11; ++c) acc[c] = 0.f;
        if (lane < P * P) {
            const int ty = lane / P, tz = lane - P * ty;
            const float uy = g.hy * (float)(oy + ty) - g.Ly2;
            const float dz = g.hz * (float)(oz + tz) - g.Lz2 - p.z;
            const float dy = uy - p.y;
            for (int tx = 0; tx < P; ++tx) {
                const float dx = g.hx * (float)(ox + tx) - g.Lx2 + g.xy * uy - p.x;
                const float w = quadw_prefac * __expf(-g.expfac * (dx * dx + dy * dy + dz * dz));
                const int r = ((bx + tx) * S + (by + ty)) * S + (bz + tz);
#pragma unroll
                for (int c = 0; c < 11; ++c) acc[c] += w * s_g[c * S3 + r];
            }
        }
#pragma unroll
        for (int c = 0; c < 11; ++c)
            for (int o = 32; o > 0; o >>= 1) acc[c] += __shfl_xor(acc[c], o, 64);
        if (lane == 0) {
            float *w = wave + 11 * (int64_t)i;
            w[0] = acc[0]; w[1] = acc[1]; w[2] = acc[2];
            // fields XX XY XZ YZ YY YX ZX ZY -> slots {XX, YX, ZX, ZY, YY, XY, XZ, YZ}
            w[3] = acc[3]; w[4] = acc[8]; w[5] = acc[9]; w[6] = acc[10];
            w[7] = acc[7]; w[8] = acc[4]; w[9] = acc[5]; w[10] = acc[6];
        }
    }
}

// The tile contraction for the usual stencil (P = 6, tile 8, S = 13) with
// compile-time sizes, the 11 fields split over two workgroups per tile
// (fields [0, 6) and [6, 11)) so that three workgroups share a CU and one's
// halo load overlaps another's arithmetic.  Same arithmetic per output as
// contract_tile_kernel.
#define CT_NF 6
#define CT_PCAP 128   // particles of a tile staged in LDS (about 45 in an 8^3 tile at volume fraction 0.45)
template <int S, int PP>
__global__ void __launch_bounds__(512) contract_tile2_kernel(SpreadGeom g, const float4 *pos, const int *orig,
                                                             const int *ostart, const int *olist, const float *grid,
                                                             int64_t M, float quadw_prefac, float *wave)
{
    constexpr int S3 = S * S * S;
    constexpr int TILE = S - PP + 1;
    extern __shared__ float s_g[];           // [CT_NF][S][S][S]
    __shared__ int s_run_beg[TILE * TILE], s_run_off[TILE * TILE + 1];
    __shared__ int s_w[3][S];                // wrapped grid index of each halo offset
    __shared__ int s_pi[CT_PCAP], s_pb[CT_PCAP];      // staged particles: index, local origin (bytes)
    __shared__ float4 s_pc[CT_PCAP];                  //   and offsets from the particle
    const int item = xcd_remap(blockIdx.x, gridDim.x);   // a slab of tiles per XCD: halos shared in its L2
    const int tile_id = item >> 1, grp = item & 1;
    const int f0 = grp ? CT_NF : 0, nf = grp ? 11 - CT_NF : CT_NF;
    const int tz_i = tile_id % g.tnz, ty_i = (tile_id / g.tnz) % g.tny, tx_i = tile_id / (g.tnz * g.tny);
    const int x0 = tx_i * TILE, y0 = ty_i * TILE, z0 = tz_i * TILE;
    const int lt = threadIdx.x;
    const int xe = min(x0 + TILE, g.Nx), ye = min(y0 + TILE, g.Ny), z1 = min(z0 + TILE, g.Nz);
    const int ncx = xe - x0, ncy = ye - y0;
    if (lt < ncx * ncy) {
        const int x = x0 + lt / ncy, y = y0 + lt % ncy;
        const int cb = (x * g.Ny + y) * g.Nz;
        const int b0 = ostart[cb + z0], b1 = ostart[cb + z1];
        s_run_beg[lt] = b0;
        s_run_off[lt + 1] = b1 - b0;
    }
    if (lt < 3 * S) {
        const int ax = lt / S, o = lt - ax * S;
        const int n = ax == 0 ? g.Nx : (ax == 1 ? g.Ny : g.Nz);
        int v = (ax == 0 ? x0 : (ax == 1 ? y0 : z0)) + o;
        v -= (v >= n) ? n : 0;
        v -= (v >= n) ? n : 0;
        s_w[ax][o] = v;
    }
    __syncthreads();
    // grid block of this workgroup's fields, wrapped: the loads of a thread are issued in
    // batches (the launch has 512 threads), so their latencies overlap
    {
        constexpr int NT = 512, ITER = (CT_NF * S3 + NT - 1) / NT, BATCH = 13;
#pragma unroll
        for (int b0 = 0; b0 < ITER; b0 += BATCH) {
            float v[BATCH];
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                const int idx = lt + (b0 + u) * NT;
                const int c = idx / S3, r = idx - c * S3;
                const int ax = r / (S * S), ay = (r / S) % S, az = r % S;
                const bool ok = b0 + u < ITER && idx < nf * S3;
                const int64_t off = ok ? (f0 + c) * M + ((int64_t)s_w[0][ax] * g.Ny + s_w[1][ay]) * g.Nz + s_w[2][az]
                                       : 0;
                v[u] = grid[off];
            }
#pragma unroll
            for (int u = 0; u < BATCH; ++u) {
                const int idx = lt + (b0 + u) * NT;
                if (b0 + u < ITER && idx < nf * S3) s_g[idx] = v[u];
            }
        }
    }
    if (lt == 0) {
        s_run_off[0] = 0;
        for (int q = 0; q < ncx * ncy; ++q) s_run_off[q + 1] += s_run_off[q];
    }
    __syncthreads();
    const int nruns = ncx * ncy;
    const int total = s_run_off[nruns];
    const int wv = lt >> 6, lane = lt & 63, nw = blockDim.x >> 6;
    // fields XX XY XZ YZ YY YX ZX ZY (grid order 3..10) -> slots {XX, YX, ZX, ZY, YY, XY, XZ, YZ}
    constexpr int slot_of[11] = {0, 1, 2, 3, 8, 9, 10, 7, 4, 5, 6};
    // olist entry of the tile's q-th particle (runs are the tile's (x, y) cell columns)
    auto entry = [&](int q) {
        int lo = 0, hi = nruns - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_run_off[mid] <= q) lo = mid; else hi = mid - 1;
        }
        return s_run_beg[lo] + (q - s_run_off[lo]);
    };
    // the tile's particles staged in LDS by one thread each, so the contraction loop below issues no
    // global load (a tile with more than CT_PCAP particles reads the rest from global memory): the
    // index, the stencil origin in the block's local frame (packed bytes) and the offsets
    // (h o - L/2 - p) of the stencil's first node from the particle, y also without p_y
    auto stage = [&](int i, int &bpk, float4 &c) {
        const int ox = orig[3 * i], oy = orig[3 * i + 1], oz = orig[3 * i + 2];
        const float4 p = pos[i];
        const int bx = (((ox % g.Nx) + g.Nx) % g.Nx) - x0;
        const int by = (((oy % g.Ny) + g.Ny) % g.Ny) - y0;
        const int bz = (((oz % g.Nz) + g.Nz) % g.Nz) - z0;
        bpk = bx | (by << 8) | (bz << 16);
        const float cy = g.hy * (float)oy - g.Ly2;
        c = make_float4(g.hx * (float)ox - g.Lx2 - p.x, cy, cy - p.y, g.hz * (float)oz - g.Lz2 - p.z);
    };
    const int nst = min(total, CT_PCAP);
    if (lt < nst) {
        const int i = olist[entry(lt)];
        int bpk;
        float4 c;
        stage(i, bpk, c);
        s_pi[lt] = i;
        s_pb[lt] = bpk;
        s_pc[lt] = c;
    }
    __syncthreads();
    // this lane's stencil column (ty, tz) and its offsets
    const int ty = lane / PP, tz = lane - PP * (lane / PP);
    const float aty = g.hy * (float)ty, atz = g.hz * (float)tz;
    const float gq = __expf(-2.0f * g.expfac * g.hx * g.hx);
    for (int q = wv; q < total; q += nw) {
        int i, bpk;
        float4 c;
        if (q < CT_PCAP) {
            i = s_pi[q];
            bpk = s_pb[q];
            c = s_pc[q];
        } else {
            i = olist[entry(q)];
            stage(i, bpk, c);
        }
        const int bx = bpk & 0xff, by = (bpk >> 8) & 0xff, bz = bpk >> 16;
        float acc[CT_NF];
#pragma unroll
        for (int c2 = 0; c2 < CT_NF; ++c2) acc[c2] = 0.f;
        if (lane < PP * PP) {
            const float uy = c.y + aty, dy = c.z + aty, dz = c.w + atz;
            const float dyz = dy * dy + dz * dz;
            const float dx0 = c.x + g.xy * uy;
            const int rb = (bx * S + by + ty) * S + bz + tz;
            // the weights along x by the fast-Gaussian-gridding recurrence (Greengard and Lee,
            // SIAM Rev. 2004): w(tx + 1) = w(tx) r(tx), r(tx + 1) = r(tx) q with
            // r(0) = exp(-a hx (2 dx0 + hx)) and q = exp(-2 a hx^2): two exponentials per lane
            float w = __expf(-g.expfac * (dx0 * dx0 + dyz));
            float r = __expf(-g.expfac * g.hx * (2.0f * dx0 + g.hx));
#pragma unroll
            for (int tx = 0; tx < PP; ++tx) {
                const int ro = rb + tx * S * S;
#pragma unroll
                for (int c2 = 0; c2 < CT_NF; ++c2)
                    if (c2 < nf) acc[c2] += w * s_g[c2 * S3 + ro];
                w *= r;
                r *= gq;
            }
        }
        // wave sums: rows of 16 by DPP, then the three rows that hold the 36 lanes.  Lane c < nf
        // writes field c and every other lane writes field 0 again (the same value to the same
        // address): one store that every lane issues, with no branch around it
        const int ce = lane < nf ? lane : 0;
        float mine = 0.f;
        int slot = 0;
#pragma unroll
        for (int f = 0; f < CT_NF; ++f) {
            const float r = group_sum<16>(acc[f]);
            const float t = quadw_prefac *
                            (__builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, r), 0)) +
                             __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, r), 16)) +
                             __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, r), 32)));
            mine = (ce == f) ? t : mine;
            slot = (ce == f) ? (f0 ? slot_of[CT_NF + (f < 11 - CT_NF ? f : 0)] : slot_of[f]) : slot;
        }
        wave[11 * (int64_t)i + slot] = mine;
    }
}

// LDS bytes of the tile contraction, 0 when it does not fit (then the
// per-particle contraction runs); sets the launch attribute once
int k_contract_lds(fsd_ctx *ctx)
{
    const int S = ctx->tile + ctx->pse.P - 1;
    if (S == 13 && ctx->pse.P == 6) {
        const size_t b2 = sizeof(float) * CT_NF * 13 * 13 * 13;
        if (hipFuncSetAttribute((const void *)contract_tile2_kernel<13, 6>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)b2) == hipSuccess) {
            ctx->contract_split = true;
            return (int)b2;
        }
    }
    const size_t bytes = sizeof(float) * 11 * (size_t)S * S * S;
    int dev = 0, maxlds = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return 0;
    if (ctx->tile * ctx->tile > 64 || bytes + 1024 > (size_t)maxlds) return 0;   // 64 origin columns per tile
    if (hipFuncSetAttribute((const void *)contract_tile_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)bytes) != hipSuccess)
        return 0;
    return (int)bytes;
}

// ------------------------------------------------------------ real space
struct RealParams {
    BoxF b;
    float rcut, dr, self_x, self_y;
    int ewald_n;
    float x_scale, inv_dr;   // ewald_n / (rcut - dr), 1 / dr
};

// Mobility_RealSpace_kernel (Mobility.cu:1046-1226) + LinearCombination/Add4
// + Mobility_D2WE_kernel + Saddle_MakeGeneralizedU_kernel (+ the saddle's
// x_U term), G lanes per particle (4, 8 or 16, the autotuner's choice; 16 by default).
// Particles are taken in cell order (`order`, the neighbour list's cell-sorted
// particles) with each XCD given one contiguous range of workgroups, so the
// neighbours gathered by concurrent particles share that XCD's L2.
template <int G>
__global__ void __launch_bounds__(256) realspace_kernel(RealParams rp, const float4 *pos, const float4 *rec,
                                                        const float4 *tab, const int *nneigh, const int *nlist,
                                                        int maxn, int N, const float *wave, const float *addU,
                                                        int use_real, float *gU, const int *order)
{
    const int gid = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int slot = gid / G, lane = gid % G;
    const bool active = slot < N;
    const int i = (active && order) ? order[slot] : slot;
    float u[3] = {0.f, 0.f, 0.f};
    float D[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (active && use_real) {
        const float4 pi = pos[i];
        const float mind2 = rp.dr * rp.dr, maxd2 = rp.rcut * rp.rcut;
        const int nn = nneigh[i];
        for (int k = lane; k < nn; k += G) {
            const int j = nlist[(int64_t)i * maxn + k];
            const float4 pj = pos[j];
            float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            box_min_image(rp.b, dx, dy, dz);
            const float d2 = dist2_nc(dx, dy, dz);
            if (!(d2 < maxd2 && d2 >= mind2)) continue;
            const float dist = sqrtf(d2);
            // table position with the divisions by constants done once on the host (x_scale =
            // ewald_n / (rcut - dr), inv_dr = 1 / dr)
            const float xr = (dist - rp.dr) * rp.x_scale;
            int r_ind = (int)floorf(xr);
            r_ind = r_ind < 0 ? 0 : (r_ind > rp.ewald_n - 1 ? rp.ewald_n - 1 : r_ind);
            const float fac = dist * rp.inv_dr - (float)r_ind - 1.0f;
            const float4 m1 = tab[2 * r_ind], p1 = tab[2 * r_ind + 2];
            const float4 m2 = tab[2 * r_ind + 1], p2 = tab[2 * r_ind + 3];
            const float f1 = m1.x + (p1.x - m1.x) * fac;
            const float f2 = m1.y + (p1.y - m1.y) * fac;
            const float g1 = -(m1.z + (p1.z - m1.z) * fac);
            const float g2 = -(m1.w + (p1.w - m1.w) * fac);
            const float h1 = m2.x + (p2.x - m2.x) * fac;
            const float h2 = m2.y + (p2.y - m2.y) * fac;
            const float h3 = m2.z + (p2.z - m2.z) * fac;
            const float idist = 1.0f / dist;
            const float R[3] = {dx * idist, dy * idist, dz * idist};
            const float4 r0 = rec[3 * (int64_t)j], r1 = rec[3 * (int64_t)j + 1], r2 = rec[3 * (int64_t)j + 2];
            const float Fj[3] = {r0.x, r0.y, r0.z};
            const float C[8] = {r0.w, r1.x, r1.y, r1.z, r1.w, r2.x, r2.y, r2.z};
            const float rdotf = R[0] * Fj[0] + R[1] * Fj[1] + R[2] * Fj[2];
            const float Cm[3][3] = {{C[0], C[1], C[2]}, {C[5], C[4], C[3]}, {C[6], C[7], -(C[0] + C[4])}};
            float Cdr[3], rdC[3];
            for (int a = 0; a < 3; ++a) {
                Cdr[a] = Cm[a][0] * R[0] + Cm[a][1] * R[1] + Cm[a][2] * R[2];
                rdC[a] = R[0] * Cm[0][a] + R[1] * Cm[1][a] + R[2] * Cm[2][a];
            }
            const float rCr = R[0] * Cdr[0] + R[1] * Cdr[1] + R[2] * Cdr[2];
            // u_a += f1 F_a + (f2 - f1)(r.F) R_a + g1 (Cdr_a - rCr R_a) + g2 (rdC_a - 4 rCr R_a)
            const float ku = (f2 - f1) * rdotf - (g1 + 4.f * g2) * rCr;
            for (int a = 0; a < 3; ++a) u[a] += f1 * Fj[a] + g1 * Cdr[a] + g2 * rdC[a] + ku * R[a];
            // D_ab += -g1 (R_a F_b - (r.F) R_a R_b) - g2 (d_ab (r.F) + F_a R_b - 4 (r.F) R_a R_b)
            //       + h1 (C_ab - 4 C_ba) + h2 (R_a Cdr_b - rCr R_a R_b)
            //       + h3 (d_ab rCr + Cdr_a R_b + R_a rdC_b + rdC_a R_b - 6 rCr R_a R_b - C_ba),
            // gathered as X_a R_b + R_a Y_b + d_ab K2 + h1 C_ab + K3 C_ba
            const float K1 = (g1 + 4.f * g2) * rdotf - (h2 + 6.f * h3) * rCr;
            const float K2 = h3 * rCr - g2 * rdotf, K3 = -(4.f * h1 + h3);
            float X[3], Y[3];
            for (int a = 0; a < 3; ++a) {
                X[a] = K1 * R[a] - g2 * Fj[a] + h3 * (Cdr[a] + rdC[a]);
                Y[a] = -g1 * Fj[a] + h2 * Cdr[a] + h3 * rdC[a];
            }
            const int A[8] = {0, 0, 0, 1, 1, 1, 2, 2};
            const int B[8] = {0, 1, 2, 2, 1, 0, 0, 1};
            const int Cs[8] = {0, 5, 6, 7, 4, 1, 2, 3};
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int a = A[c], bb = B[c];
                float d = X[a] * R[bb] + R[a] * Y[bb] + h1 * C[c] + K3 * C[Cs[c]];
                if (a == bb) d += K2;
                D[c] += d;
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) u[a] = group_sum<G>(u[a]);
#pragma unroll
    for (int c = 0; c < 8; ++c) D[c] = group_sum<G>(D[c]);
    if (!active || lane != 0) return;
    const float4 r0 = rec[3 * (int64_t)i], r1 = rec[3 * (int64_t)i + 1], r2 = rec[3 * (int64_t)i + 2];
    if (use_real) {
        const float C[8] = {r0.w, r1.x, r1.y, r1.z, r1.w, r2.x, r2.y, r2.z};
        u[0] += rp.self_x * r0.x; u[1] += rp.self_x * r0.y; u[2] += rp.self_x * r0.z;
        D[0] += rp.self_y * (C[0] - 4.f * C[0]); D[1] += rp.self_y * (C[1] - 4.f * C[5]);
        D[2] += rp.self_y * (C[2] - 4.f * C[6]); D[3] += rp.self_y * (C[3] - 4.f * C[7]);
        D[4] += rp.self_y * (C[4] - 4.f * C[4]); D[5] += rp.self_y * (C[5] - 4.f * C[1]);
        D[6] += rp.self_y * (C[6] - 4.f * C[2]); D[7] += rp.self_y * (C[7] - 4.f * C[3]);
    }
    if (wave) {
        const float *w = wave + 11 * (int64_t)i;
        u[0] += w[0]; u[1] += w[1]; u[2] += w[2];
        for (int c = 0; c < 8; ++c) D[c] += w[3 + c];
    }
    // D2WE (Helper_Mobility.cu:235-283)
    const float W0 = 0.5f * (D[3] - D[7]), W1 = 0.5f * (D[6] - D[2]), W2 = 0.5f * (D[1] - D[5]);
    const float e0 = D[0], e1 = 0.5f * (D[1] + D[5]), e2 = 0.5f * (D[2] + D[6]), e3 = 0.5f * (D[3] + D[7]), e4 = D[4];
    float *o6 = gU + 6 * (int64_t)i;
    float add[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (addU)
        for (int a = 0; a < 6; ++a) add[a] = addU[6 * (int64_t)i + a];
    o6[0] = u[0] + add[0]; o6[1] = u[1] + add[1]; o6[2] = u[2] + add[2];
    o6[3] = W0 + add[3]; o6[4] = W1 + add[4]; o6[5] = W2 + add[5];
    float *o5 = gU + 6 * (int64_t)N + 5 * (int64_t)i;
    o5[0] = 2.f * e0 + e4; o5[1] = 2.f * e1; o5[2] = 2.f * e2; o5[3] = 2.f * e3; o5[4] = 2.f * e4 + e0;
}

// inverse transform of the 11 half spectra in d_spec and contraction of the
// grids at the particles into d_wave (u and the velocity gradient); with
// inverse_yz only the (y, z) part of the inverse remains (x has been done)
int k_wave_contract(fsd_ctx *ctx, const float4 *pos, bool inverse_yz)
{
    const int N = ctx->N;
    const fsd_pse_info &ps = ctx->pse;
    if (!ctx->spread_valid || ctx->spread_pos != (const float *)pos) {
        const int rc = k_setup_spread(ctx, pos);
        if (rc) return rc;
    }
    SpreadGeom g;
    g.Nx = ps.Nx; g.Ny = ps.Ny; g.Nz = ps.Nz; g.P = ps.P; g.tile = ctx->tile;
    g.tnx = ctx->tiles[0]; g.tny = ctx->tiles[1]; g.tnz = ctx->tiles[2];
    g.hx = ps.gridh_x; g.hy = ps.gridh_y; g.hz = ps.gridh_z;
    g.Lx2 = ctx->box.Lx / 2.0f; g.Ly2 = ctx->box.Ly / 2.0f; g.Lz2 = ctx->box.Lz / 2.0f; g.xy = ctx->box.xy;
    g.prefac = ps.prefac; g.expfac = ps.expfac;
    hipEvent_t ev;
    prof_begin(ctx, PROF_FFT, &ev);
    {
        void *in[1] = {ctx->d_spec};
        void *out[1] = {ctx->d_grid};
        if (rocfft_execute(inverse_yz ? ctx->plan_inv2 : ctx->plan_inv, in, out, ctx->info_inv) !=
            rocfft_status_success) {
            ctx->err = "rocfft inverse failed";
            return FSD_E_HIP;
        }
    }
    prof_end(ctx, PROF_FFT, ev);
    prof_begin(ctx, PROF_CONTRACT, &ev);
    const bool tiled = ctx->tune.contract == 0 && ctx->contract_lds;
    if (tiled && ctx->contract_split)
        hipLaunchKernelGGL((contract_tile2_kernel<13, 6>), dim3(2 * ctx->ntiles), dim3(512), ctx->contract_lds,
                           ctx->stream, g, pos, ctx->d_porig, ctx->d_ostart, ctx->d_olist, ctx->d_grid, (int64_t)ctx->M,
                           ps.quadW * ps.prefac, ctx->d_wave);
    else if (tiled)
        hipLaunchKernelGGL(contract_tile_kernel, dim3(ctx->ntiles), dim3(512), ctx->contract_lds, ctx->stream, g, pos,
                           ctx->d_porig, ctx->d_ostart, ctx->d_olist, ctx->d_grid, (int64_t)ctx->M,
                           ps.quadW * ps.prefac, ctx->d_wave);
    else
        hipLaunchKernelGGL(contract_kernel, dim3(nblk((int64_t)N * 64, 256)), dim3(256), 0, ctx->stream, g, pos,
                           ctx->d_porig, ctx->d_grid, (int64_t)ctx->M, N, ps.quadW * ps.prefac, ctx->d_wave);
    FSD_HIP(hipGetLastError());
    prof_end(ctx, PROF_CONTRACT, ev);
    return 0;
}

// ------------------------------------------------------------ driver
// gU = M gF (+ addU on the first 6N rows); parts: 1 real, 2 wave, 3 both
int k_mobility(fsd_ctx *ctx, const float4 *pos, const float *gF, float *gU, int parts, const float *addU)
{
    const int N = ctx->N;
    const fsd_pse_info &ps = ctx->pse;
    float4 *rec = ctx->d_rec;
    hipLaunchKernelGGL(fc_records_kernel, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, gF, N, rec);
    FSD_HIP(hipGetLastError());
    SpreadGeom g;
    g.Nx = ps.Nx; g.Ny = ps.Ny; g.Nz = ps.Nz; g.P = ps.P; g.tile = ctx->tile;
    g.tnx = ctx->tiles[0]; g.tny = ctx->tiles[1]; g.tnz = ctx->tiles[2];
    g.hx = ps.gridh_x; g.hy = ps.gridh_y; g.hz = ps.gridh_z;
    g.Lx2 = ctx->box.Lx / 2.0f; g.Ly2 = ctx->box.Ly / 2.0f; g.Lz2 = ctx->box.Lz / 2.0f; g.xy = ctx->box.xy;
    g.prefac = ps.prefac; g.expfac = ps.expfac;
    const float *wave = nullptr;
    if (parts & 2) {
        if (!ctx->spread_valid || ctx->spread_pos != (const float *)pos) {
            int rc = k_setup_spread(ctx, pos);
            if (rc) return rc;
        }
        hipEvent_t ev;
        prof_begin(ctx, PROF_SPREAD, &ev);
        if (ctx->tile == 8 && ctx->pse.P == 6 && ctx->tune.spread == 0)
            hipLaunchKernelGGL((spread_tile2_kernel<8, 6>), dim3(ctx->ntiles), dim3(512), 0, ctx->stream, g, pos,
                               ctx->d_porig, rec, ctx->d_ostart, ctx->d_olist, ctx->d_grid, (int64_t)ctx->M);
        else
            hipLaunchKernelGGL(spread_kernel, dim3(ctx->ntiles), dim3(ctx->tile * ctx->tile * ctx->tile), 0,
                               ctx->stream, g, pos, ctx->d_porig, rec, ctx->d_ostart, ctx->d_olist, ctx->d_grid,
                               (int64_t)ctx->M);
        FSD_HIP(hipGetLastError());
        prof_end(ctx, PROF_SPREAD, ev);
        const bool fx = ctx->tune.xpass == 0 && ctx->fx_B > 0 && ctx->plan_fwd2 && ctx->plan_inv2;
        prof_begin(ctx, PROF_FFT, &ev);
        {
            void *in[1] = {ctx->d_grid};
            void *out[1] = {ctx->d_spec};
            if (rocfft_execute(fx ? ctx->plan_fwd2 : ctx->plan_fwd, in, out, ctx->info_fwd) != rocfft_status_success) {
                ctx->err = "rocfft forward failed";
                return FSD_E_HIP;
            }
        }
        prof_end(ctx, PROF_FFT, ev);
        prof_begin(ctx, PROF_GREEN, &ev);
        GreenGeom gg{ps.Nx, ps.Ny, ps.Nz, ctx->box, ctx->p.xi, ps.eta, k_twopi(ctx)};
        if (fx) {   // x transforms and the Green's function in one pass
            if (!launch_fft_x_green(ctx, gg)) { ctx->err = "no x-transform kernel for this grid"; return FSD_E_PARAM; }
        } else {
            hipLaunchKernelGGL(green_kernel, dim3(nblk(ctx->Mh, 256)), dim3(256), 0, ctx->stream, ctx->d_spec,
                               ctx->d_gridk, (int64_t)ctx->Mh, gg);
        }
        FSD_HIP(hipGetLastError());
        prof_end(ctx, PROF_GREEN, ev);
        const int rcw = k_wave_contract(ctx, pos, fx);
        if (rcw) return rcw;
        wave = ctx->d_wave;
    }
    RealParams rp;
    rp.b = ctx->box; rp.rcut = ps.rcut; rp.dr = 0.001f; rp.self_x = ps.self_x; rp.self_y = ps.self_y;
    rp.ewald_n = ps.ewald_n;
    rp.x_scale = (float)((double)ps.ewald_n / ((double)ps.rcut - (double)rp.dr));
    rp.inv_dr = 1.0f / rp.dr;
    // the processing order is the neighbour list's cell list (identity without a list, wave part
    // only, or for a box of fewer than three cells per side, whose list is built over all pairs)
    if (!ctx->nlist_valid && (parts & 1)) { ctx->err = "neighbour list not built"; return FSD_E_STATE; }
    const int *order = (ctx->nlist_valid && ctx->use_cells) ? ctx->d_cell_list : nullptr;
    hipEvent_t ev;
    prof_begin(ctx, PROF_REAL, &ev);
    const int G = ctx->tune.real_g, blk = ctx->tune.real_blk;
    const dim3 rg(nblk((int64_t)N * G, blk)), rb(blk);
    const float4 *tab = reinterpret_cast<const float4 *>(ctx->d_ewald);
    const int ur = (parts & 1) ? 1 : 0;
    if (G == 4)
        hipLaunchKernelGGL(realspace_kernel<4>, rg, rb, 0, ctx->stream, rp, pos, rec, tab, ctx->d_nneigh, ctx->d_nlist,
                           ctx->maxn, N, wave, addU, ur, gU, order);
    else if (G == 8)
        hipLaunchKernelGGL(realspace_kernel<8>, rg, rb, 0, ctx->stream, rp, pos, rec, tab, ctx->d_nneigh, ctx->d_nlist,
                           ctx->maxn, N, wave, addU, ur, gU, order);
    else
        hipLaunchKernelGGL(realspace_kernel<16>, rg, rb, 0, ctx->stream, rp, pos, rec, tab, ctx->d_nneigh,
                           ctx->d_nlist, ctx->maxn, N, wave, addU, ur, gU, order);
    FSD_HIP(hipGetLastError());
    prof_end(ctx, PROF_REAL, ev);
    return 0;
}

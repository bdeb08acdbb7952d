// Neighbour list (cell list), lubrication operators and the interparticle
// force law.  All kernels are HBM/L2 gather-bound; the per-step lubrication
// pair records turn every R_FU apply of the GMRES loop into a pure block
// gather with no table lookups.
#include "fsd_internal.h"

// ---------------------------------------------------------------- scans
// single-workgroup exclusive scan (n up to a few 10^5); out[n] = total
__global__ void __launch_bounds__(1024) scan_excl_kernel(const int *in, int *out, int n)
{
    __shared__ int warp_sums[16];
    __shared__ int carry;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 1024) {
        int i = base + tid;
        int v = (i < n) ? in[i] : 0;
        // inclusive scan within the wavefront
        int x = v;
        for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) warp_sums[wid] = x;
        __syncthreads();
        if (wid == 0) {
            int s = (lane < 16) ? warp_sums[lane] : 0;
            for (int o = 1; o < 16; o <<= 1) {
                int y = __shfl_up(s, o, 64);
                if (lane >= o) s += y;
            }
            if (lane < 16) warp_sums[lane] = s;
        }
        __syncthreads();
        int prefix = carry + ((wid > 0) ? warp_sums[wid - 1] : 0);
        if (i < n) out[i] = prefix + x - v;
        __syncthreads();
        if (tid == 1023) carry = prefix + x;
        __syncthreads();
    }
    if (tid == 0) out[n] = carry;
}

// total of n ints into *out (single workgroup)
__global__ void __launch_bounds__(1024) sum_int_kernel(const int *in, int n, int *out)
{
    __shared__ int part[16];
    int s = 0;
    for (int i = threadIdx.x; i < n; i += 1024) s += in[i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < 16; ++w) t += part[w];
        *out = t;
    }
}

int scan_excl(fsd_ctx *ctx, const int *in, int *out, int n)
{
    hipLaunchKernelGGL(scan_excl_kernel, dim3(1), dim3(1024), 0, ctx->stream, in, out, n);
    FSD_HIP(hipGetLastError());
    return 0;
}

// multi-workgroup exclusive scan for long arrays: each workgroup scans
// SCAN_CHUNK elements (4 per thread) and records its total; the totals are
// scanned by one workgroup; a third pass adds each chunk's offset.
#define SCAN_CHUNK 4096
__global__ void __launch_bounds__(1024) scan_chunks_kernel(const int *in, int *out, int n, int *part)
{
    __shared__ int wsum[16];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK + 4 * tid;
    int v[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t i = base + k;
        v[k] = (i < n) ? in[i] : 0;
        s += v[k];
    }
    int x = s;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    if (wid == 0) {
        int t = (lane < 16) ? wsum[lane] : 0;
        for (int o = 1; o < 16; o <<= 1) {
            const int y = __shfl_up(t, o, 64);
            if (lane >= o) t += y;
        }
        if (lane < 16) wsum[lane] = t;
    }
    __syncthreads();
    int pre = ((wid > 0) ? wsum[wid - 1] : 0) + x - s;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t i = base + k;
        if (i < n) out[i] = pre;
        pre += v[k];
    }
    if (tid == 0) part[blockIdx.x] = wsum[15];
}

__global__ void scan_add_kernel(int *out, int n, const int *part_scan, int nchunks)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += part_scan[i / SCAN_CHUNK];
    if (i == 0) out[n] = part_scan[nchunks];
}

int scan_excl_chunks(int n) { return (n + SCAN_CHUNK - 1) / SCAN_CHUNK; }

// out[0..n] = exclusive scan of in[0..n); tmp holds 2 (chunks + 1) ints
int scan_excl_large(fsd_ctx *ctx, const int *in, int *out, int n, int *tmp)
{
    const int nc = scan_excl_chunks(n);
    if (nc <= 1) return scan_excl(ctx, in, out, n);
    int *part = tmp, *part_scan = tmp + nc + 1;
    hipLaunchKernelGGL(scan_chunks_kernel, dim3(nc), dim3(1024), 0, ctx->stream, in, out, n, part);
    hipLaunchKernelGGL(scan_excl_kernel, dim3(1), dim3(1024), 0, ctx->stream, part, part_scan, nc);
    hipLaunchKernelGGL(scan_add_kernel, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, out, n, part_scan, nc);
    FSD_HIP(hipGetLastError());
    return 0;
}

// ----------------------------------------------------------- cell list
__global__ void cell_assign_kernel(const float4 *pos, int N, BoxF b, int cx, int cy, int cz, int *cell_of, int *count)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float4 p = pos[i];
    float fx, fy, fz;
    box_make_fraction(b, p.x, p.y, p.z, fx, fy, fz);
    int ix = (int)floorf(fx * cx), iy = (int)floorf(fy * cy), iz = (int)floorf(fz * cz);
    ix = ix < 0 ? 0 : (ix >= cx ? cx - 1 : ix);
    iy = iy < 0 ? 0 : (iy >= cy ? cy - 1 : iy);
    iz = iz < 0 ? 0 : (iz >= cz ? cz - 1 : iz);
    int c = (ix * cy + iy) * cz + iz;
    cell_of[i] = c;
    atomicAdd(&count[c], 1);
}

__global__ void cell_scatter_kernel(const int *cell_of, int N, const int *start, int *fill, int *list)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int c = cell_of[i];
    int slot = atomicAdd(&fill[c], 1);
    list[start[c] + slot] = i;
}

// deterministic order inside a cell: ascending particle index
__global__ void cell_sort_kernel(const int *start, int ncells, int *list)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    int s = start[c], e = start[c + 1];
    for (int a = s + 1; a < e; ++a) {
        int key = list[a], q = a - 1;
        while (q >= s && list[q] > key) { list[q + 1] = list[q]; --q; }
        list[q + 1] = key;
    }
}

__global__ void nlist_cells_kernel(const float4 *pos, int N, BoxF b, int cx, int cy, int cz, const int *cell_of,
                                   const int *start, const int *list, float rl2, int maxn, int *nneigh, int *nlist,
                                   int *overflow)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float4 pi = pos[i];
    int c = cell_of[i];
    int ix = c / (cy * cz), iy = (c / cz) % cy, iz = c % cz;
    int cnt = 0;
    for (int dx = -1; dx <= 1; ++dx) {
        int jx = ix + dx; jx = jx < 0 ? jx + cx : (jx >= cx ? jx - cx : jx);
        for (int dy = -1; dy <= 1; ++dy) {
            int jy = iy + dy; jy = jy < 0 ? jy + cy : (jy >= cy ? jy - cy : jy);
            for (int dz = -1; dz <= 1; ++dz) {
                int jz = iz + dz; jz = jz < 0 ? jz + cz : (jz >= cz ? jz - cz : jz);
                int cc = (jx * cy + jy) * cz + jz;
                for (int q = start[cc]; q < start[cc + 1]; ++q) {
                    int j = list[q];
                    if (j == i) continue;
                    float4 pj = pos[j];
                    float ddx = pj.x - pi.x, ddy = pj.y - pi.y, ddz = pj.z - pi.z;
                    box_min_image(b, ddx, ddy, ddz);
                    if (dist2_nc(ddx, ddy, ddz) < rl2) {
                        if (cnt < maxn) nlist[(int64_t)i * maxn + cnt] = j;
                        ++cnt;
                    }
                }
            }
        }
    }
    if (cnt > maxn) atomicMax(overflow, cnt);
    nneigh[i] = cnt < maxn ? cnt : maxn;
}

// small boxes (fewer than 3 cells along an axis): all pairs
__global__ void nlist_allpairs_kernel(const float4 *pos, int N, BoxF b, float rl2, int maxn, int *nneigh, int *nlist,
                                      int *overflow)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float4 pi = pos[i];
    int cnt = 0;
    for (int j = 0; j < N; ++j) {
        if (j == i) continue;
        float4 pj = pos[j];
        float ddx = pj.x - pi.x, ddy = pj.y - pi.y, ddz = pj.z - pi.z;
        box_min_image(b, ddx, ddy, ddz);
        if (dist2_nc(ddx, ddy, ddz) < rl2) {
            if (cnt < maxn) nlist[(int64_t)i * maxn + cnt] = j;
            ++cnt;
        }
    }
    if (cnt > maxn) atomicMax(overflow, cnt);
    nneigh[i] = cnt < maxn ? cnt : maxn;
}

// NeighborList::distanceCheck of HOOMD: some particle has moved at least r_buff / 2 (minimum
// image) since the list was built
__global__ void max_disp_kernel(const float4 *pos, const float4 *last, int N, BoxF b, float lim2, int *flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float4 p = pos[i], q = last[i];
    float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    box_min_image(b, dx, dy, dz);
    if (dist2_nc(dx, dy, dz) >= lim2) *flag = 1;
}

// the list is kept while no particle has moved r_buff / 2 since it was built (the buffer then
// still holds every pair within rcut), rebuilt otherwise or when the box changed
int k_update_nlist(fsd_ctx *ctx, const float4 *pos, bool *rebuilt)
{
    *rebuilt = false;
    if (ctx->nlist_valid) {
        const float half = 0.5f * ctx->p.r_buff;
        FSD_HIP(hipMemsetAsync(ctx->d_flags + 6, 0, sizeof(int), ctx->stream));
        hipLaunchKernelGGL(max_disp_kernel, dim3(nblk(ctx->N, 256)), dim3(256), 0, ctx->stream, pos, ctx->d_pos_last,
                           ctx->N, ctx->box, half * half, ctx->d_flags + 6);
        FSD_HIP(hipGetLastError());
        FSD_HIP(hipMemcpyAsync(ctx->h_flags + 6, ctx->d_flags + 6, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        FSD_HIP(hipStreamSynchronize(ctx->stream));
        if (!ctx->h_flags[6]) return 0;
    }
    *rebuilt = true;
    return k_build_nlist(ctx, pos);
}

int k_build_nlist(fsd_ctx *ctx, const float4 *pos)
{
    const int N = ctx->N;
    const float rl = ctx->pse.rcut + ctx->p.r_buff;
    const float rl2 = rl * rl;
    hipEvent_t ev;
    prof_begin(ctx, PROF_NLIST, &ev);
    for (int attempt = 0; attempt < 2; ++attempt) {
        FSD_HIP(hipMemsetAsync(ctx->d_flags, 0, 8 * sizeof(int), ctx->stream));
        if (ctx->use_cells) {
            FSD_HIP(hipMemsetAsync(ctx->d_cell_count, 0, sizeof(int) * ctx->ncells, ctx->stream));
            hipLaunchKernelGGL(cell_assign_kernel, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, pos, N, ctx->box,
                               ctx->ncell[0], ctx->ncell[1], ctx->ncell[2], ctx->d_cell_of, ctx->d_cell_count);
            if (scan_excl(ctx, ctx->d_cell_count, ctx->d_cell_start, ctx->ncells)) return FSD_E_HIP;
            FSD_HIP(hipMemsetAsync(ctx->d_cell_count, 0, sizeof(int) * ctx->ncells, ctx->stream));
            hipLaunchKernelGGL(cell_scatter_kernel, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, ctx->d_cell_of, N,
                               ctx->d_cell_start, ctx->d_cell_count, ctx->d_cell_list);
            hipLaunchKernelGGL(cell_sort_kernel, dim3(nblk(ctx->ncells, 128)), dim3(128), 0, ctx->stream,
                               ctx->d_cell_start, ctx->ncells, ctx->d_cell_list);
            hipLaunchKernelGGL(nlist_cells_kernel, dim3(nblk(N, 128)), dim3(128), 0, ctx->stream, pos, N, ctx->box,
                               ctx->ncell[0], ctx->ncell[1], ctx->ncell[2], ctx->d_cell_of, ctx->d_cell_start,
                               ctx->d_cell_list, rl2, ctx->maxn, ctx->d_nneigh, ctx->d_nlist, ctx->d_flags);
        } else {
            hipLaunchKernelGGL(nlist_allpairs_kernel, dim3(nblk(N, 128)), dim3(128), 0, ctx->stream, pos, N, ctx->box,
                               rl2, ctx->maxn, ctx->d_nneigh, ctx->d_nlist, ctx->d_flags);
        }
        FSD_HIP(hipGetLastError());
        hipLaunchKernelGGL(sum_int_kernel, dim3(1), dim3(1024), 0, ctx->stream, ctx->d_nneigh, N, ctx->d_flags + 4);
        FSD_HIP(hipMemcpyAsync(ctx->h_flags, ctx->d_flags, 8 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        FSD_HIP(hipStreamSynchronize(ctx->stream));
        int need = ctx->h_flags[0];
        ctx->nlist_total = ctx->h_flags[4];
        if (need == 0) break;
        if (attempt == 1) { ctx->err = "neighbour list overflow after regrow"; return FSD_E_CAPACITY; }
        // grow and rebuild
        int newmax = ((need + 32) / 32) * 32 + 32;
        FSD_HIP(hipFree(ctx->d_nlist));
        FSD_HIP(hipMalloc(&ctx->d_nlist, sizeof(int) * (size_t)N * newmax));
        ctx->maxn = newmax;
    }
    // positions the list was built at, for the rebuild check
    FSD_HIP(hipMemcpyAsync(ctx->d_pos_last, pos, sizeof(float4) * (size_t)N, hipMemcpyDeviceToDevice, ctx->stream));
    prof_end(ctx, PROF_NLIST, ev);
    ctx->nlist_valid = true;
    ctx->lub_valid = false;
    ctx->spread_valid = false;
    ctx->precond_valid = false;
    return 0;
}

// -------------------------------------------------- lubrication records
// One thread per particle: select neighbours with 0 < r^2 < rlub^2, store
// the pair record and accumulate the particle's 6x6 self block
// (Lubrication.cu:147-296 split into j- and i-side terms).
__global__ void lub_pairs_kernel(const float4 *pos, int N, BoxF b, const int *nneigh, const int *nlist, int maxn,
                                 LubTable t, float rl2, int maxlub, int *lub_n, int *lub_j, float4 *lub_a,
                                 float4 *lub_b, float *lub_self, int *overflow)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float4 pi = pos[i];
    float S[36];
    for (int q = 0; q < 36; ++q) S[q] = 0.f;
    int cnt = 0;
    const int nn = nneigh[i];
    for (int k = 0; k < nn; ++k) {
        int j = nlist[(int64_t)i * maxn + k];
        float4 pj = pos[j];
        float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        box_min_image(b, dx, dy, dz);
        float d2 = dist2_nc(dx, dy, dz);
        if (!(d2 < rl2 && d2 > 0.f)) continue;
        float dist = sqrtf(d2);
        float v[10];
        lub_lookup(t, dist, 0, 10, 232, v);
        float rx = dx / dist, ry = dy / dist, rz = dz / dist;
        if (cnt < maxlub) {
            int64_t o = (int64_t)i * maxlub + cnt;
            lub_j[o] = j;
            lub_a[o] = make_float4(rx, ry, rz, v[1] - v[3]);
            lub_b[o] = make_float4(v[3], v[5], v[7] - v[9], v[9]);
        }
        ++cnt;
        const float r[3] = {rx, ry, rz};
        // E v = eps.r.v  (Lubrication.cu:243-254)
        const float E[3][3] = {{0.f, rz, -ry}, {-rz, 0.f, rx}, {ry, -rx, 0.f}};
        const float XA11 = v[0], YA11 = v[2], YB11 = v[4], XC11 = v[6], YC11 = v[8];
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) {
                float rr = r[a] * r[c];
                float del = (a == c) ? 1.f : 0.f;
                S[a * 6 + c] += (XA11 - YA11) * rr + YA11 * del;
                S[a * 6 + 3 + c] += -YB11 * E[a][c];
                S[(3 + a) * 6 + c] += YB11 * E[a][c];
                S[(3 + a) * 6 + 3 + c] += (XC11 - YC11) * rr + YC11 * del;
            }
    }
    if (cnt > maxlub) atomicMax(overflow, cnt);
    lub_n[i] = cnt < maxlub ? cnt : maxlub;
    for (int q = 0; q < 36; ++q) lub_self[(int64_t)i * 36 + q] = S[q];
}

int k_build_lub_pairs(fsd_ctx *ctx, const float4 *pos)
{
    const int N = ctx->N;
    if (!ctx->have_table) { ctx->err = "resistance table not set"; return FSD_E_STATE; }
    LubTable t{ctx->d_res_dist, ctx->d_res_vals, ctx->table_min, ctx->table_dr};
    const float rl2 = ctx->p.rlub * ctx->p.rlub;
    for (int attempt = 0; attempt < 2; ++attempt) {
        FSD_HIP(hipMemsetAsync(ctx->d_flags + 1, 0, sizeof(int), ctx->stream));
        hipLaunchKernelGGL(lub_pairs_kernel, dim3(nblk(N, 128)), dim3(128), 0, ctx->stream, pos, N, ctx->box,
                           ctx->d_nneigh, ctx->d_nlist, ctx->maxn, t, rl2, ctx->maxlub, ctx->d_lub_n, ctx->d_lub_j,
                           ctx->d_lub_a, ctx->d_lub_b, ctx->d_lub_self, ctx->d_flags + 1);
        FSD_HIP(hipGetLastError());
        hipLaunchKernelGGL(sum_int_kernel, dim3(1), dim3(1024), 0, ctx->stream, ctx->d_lub_n, N, ctx->d_flags + 5);
        FSD_HIP(hipMemcpyAsync(ctx->h_flags + 1, ctx->d_flags + 1, 5 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        FSD_HIP(hipStreamSynchronize(ctx->stream));
        ctx->lub_total = ctx->h_flags[5];
        int need = ctx->h_flags[1];
        if (need == 0) break;
        if (attempt == 1) { ctx->err = "lubrication pair capacity exceeded"; return FSD_E_CAPACITY; }
        int nm = ((need + 16) / 16) * 16 + 16;
        FSD_HIP(hipFree(ctx->d_lub_j)); FSD_HIP(hipFree(ctx->d_lub_a)); FSD_HIP(hipFree(ctx->d_lub_b));
        FSD_HIP(hipMalloc(&ctx->d_lub_j, sizeof(int) * (size_t)N * nm));
        FSD_HIP(hipMalloc(&ctx->d_lub_a, sizeof(float4) * (size_t)N * nm));
        FSD_HIP(hipMalloc(&ctx->d_lub_b, sizeof(float4) * (size_t)N * nm));
        ctx->maxlub = nm;
    }
    ctx->lub_valid = true;
    return 0;
}

// Pair records into caller-provided buffers (capacity maxlub per particle); the
// overflow flag must be zero on entry and is left > 0 when some particle has more
// pairs than maxlub.  No allocation: used by the allocation-free entry point.
int k_lub_pairs_into(fsd_ctx *ctx, const float4 *pos, int maxlub, int *lub_n, int *lub_j, float4 *lub_a,
                     float4 *lub_b, float *lub_self, int *overflow)
{
    if (!ctx->have_table) { ctx->err = "resistance table not set"; return FSD_E_STATE; }
    LubTable t{ctx->d_res_dist, ctx->d_res_vals, ctx->table_min, ctx->table_dr};
    const float rl2 = ctx->p.rlub * ctx->p.rlub;
    hipLaunchKernelGGL(lub_pairs_kernel, dim3(nblk(ctx->N, 128)), dim3(128), 0, ctx->stream, pos, ctx->N, ctx->box,
                       ctx->d_nneigh, ctx->d_nlist, ctx->maxn, t, rl2, maxlub, lub_n, lub_j, lub_a, lub_b, lub_self,
                       overflow);
    FSD_HIP(hipGetLastError());
    return 0;
}

// R_FU apply: G lanes per particle (8 or 16: lanes 0-5 write the rows).  out[6i+r] = sign_out * (R U)_ir
// (+ xF[6i+r] when xF != nullptr, i.e. the saddle row  x_F - R_FU x_U).
template <int G>
__global__ void __launch_bounds__(256) rfu_apply_kernel(int N, const float *U, const int *lub_n, const int *lub_j,
                                                        const float4 *lub_a, const float4 *lub_b, const float *lub_self,
                                                        int maxlub, const float *xF, float sign_out, float *out,
                                                        const int *order)
{
    // particles in cell order, a contiguous range of them per XCD (neighbour gathers share its L2)
    const int gid = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int slot = gid / G, lane = gid % G;
    const bool active = slot < N;
    const int i = (active && order) ? order[slot] : slot;
    float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (active) {
        const int n = lub_n[i];
        for (int k = lane; k < n; k += G) {
            const int64_t o = (int64_t)i * maxlub + k;
            const int j = lub_j[o];
            const float4 a = lub_a[o], bb = lub_b[o];
            const float2 u01 = *reinterpret_cast<const float2 *>(U + 6 * (int64_t)j);
            const float2 u23 = *reinterpret_cast<const float2 *>(U + 6 * (int64_t)j + 2);
            const float2 u45 = *reinterpret_cast<const float2 *>(U + 6 * (int64_t)j + 4);
            const float ux = u01.x, uy = u01.y, uz = u23.x, wx = u23.y, wy = u45.x, wz = u45.y;
            const float rx = a.x, ry = a.y, rz = a.z;
            const float dA = a.w, YA = bb.x, YB = bb.y, dC = bb.z, YC = bb.w;
            const float rdu = rx * ux + ry * uy + rz * uz;
            const float rdw = rx * wx + ry * wy + rz * wz;
            // eps.r.v
            const float eux = rz * uy - ry * uz, euy = -rz * ux + rx * uz, euz = ry * ux - rx * uy;
            const float ewx = rz * wy - ry * wz, ewy = -rz * wx + rx * wz, ewz = ry * wx - rx * wy;
            acc[0] += dA * rdu * rx + YA * ux + YB * ewx;
            acc[1] += dA * rdu * ry + YA * uy + YB * ewy;
            acc[2] += dA * rdu * rz + YA * uz + YB * ewz;
            acc[3] += YB * eux + dC * rdw * rx + YC * wx;
            acc[4] += YB * euy + dC * rdw * ry + YC * wy;
            acc[5] += YB * euz + dC * rdw * rz + YC * wz;
        }
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) acc[r] = group_sum<G>(acc[r]);
    if (active && lane < 6) {
        const float *S = lub_self + (int64_t)i * 36 + lane * 6;
        const float *ui = U + 6 * (int64_t)i;
        float s = 0.f;
        for (int c = 0; c < 6; ++c) s += S[c] * ui[c];
        float v = acc[0];
        // select acc[lane] without dynamic register indexing
        v = (lane == 1) ? acc[1] : v; v = (lane == 2) ? acc[2] : v; v = (lane == 3) ? acc[3] : v;
        v = (lane == 4) ? acc[4] : v; v = (lane == 5) ? acc[5] : v;
        float f = sign_out * (s + v);
        if (xF) f += xF[6 * (int64_t)i + lane];
        out[6 * (int64_t)i + lane] = f;
    }
}

int k_rfu_apply(fsd_ctx *ctx, const float *U, float *F, const float *xF, float sign_out)
{
    // the processing order is the neighbour list's cell list; a box of fewer than three cells per
    // side builds its list over all pairs and has no cell list (identity order)
    if (!ctx->nlist_valid) { ctx->err = "neighbour list not built"; return FSD_E_STATE; }
    const int *order = ctx->use_cells ? ctx->d_cell_list : nullptr;
    const int N = ctx->N;
    hipEvent_t ev;
    prof_begin(ctx, PROF_RFU, &ev);
    const int G = ctx->tune.rfu_g, blk = ctx->tune.rfu_blk;
    const dim3 rg(nblk((int64_t)N * G, blk)), rb(blk);
    if (G == 16)
        hipLaunchKernelGGL(rfu_apply_kernel<16>, rg, rb, 0, ctx->stream, N, U, ctx->d_lub_n, ctx->d_lub_j,
                           ctx->d_lub_a, ctx->d_lub_b, ctx->d_lub_self, ctx->maxlub, xF, sign_out, F, order);
    else
        hipLaunchKernelGGL(rfu_apply_kernel<8>, rg, rb, 0, ctx->stream, N, U, ctx->d_lub_n, ctx->d_lub_j,
                           ctx->d_lub_a, ctx->d_lub_b, ctx->d_lub_self, ctx->maxlub, xF, sign_out, F, order);
    FSD_HIP(hipGetLastError());
    prof_end(ctx, PROF_RFU, ev);
    return 0;
}

// R_FU apply from caller-provided pair records (the allocation-free entry point)
int k_rfu_apply_with(fsd_ctx *ctx, const float *U, float *F, const int *lub_n, const int *lub_j, const float4 *lub_a,
                     const float4 *lub_b, const float *lub_self, int maxlub)
{
    if (!ctx->nlist_valid) { ctx->err = "neighbour list not built"; return FSD_E_STATE; }
    const int *order = ctx->use_cells ? ctx->d_cell_list : nullptr;
    const int N = ctx->N;
    hipLaunchKernelGGL(rfu_apply_kernel<8>, dim3(nblk((int64_t)N * 8, 256)), dim3(256), 0, ctx->stream, N, U, lub_n,
                       lub_j, lub_a, lub_b, lub_self, maxlub, (const float *)nullptr, 1.0f, F, order);
    FSD_HIP(hipGetLastError());
    return 0;
}

// ----------------------------------------------------- RFE / RSU / RSE
struct EParams { float shear_rate, B2; const float *mask2; const float *ori; };

__device__ inline void strain_block(const EParams &ep, int k, float E[3][3])
{
    float e[5] = {0.f, ep.shear_rate, 0.f, 0.f, 0.f};
    if (ep.B2 != 0.f) {
        const float b2 = 0.6f * ep.B2 * ep.mask2[k];
        const float px = ep.ori[3 * k], py = ep.ori[3 * k + 1], pz = ep.ori[3 * k + 2];
        e[0] += b2 * (px * px - pz * pz); e[1] += b2 * (2.f * px * py); e[2] += b2 * (2.f * px * pz);
        e[3] += b2 * (2.f * py * pz); e[4] += b2 * (py * py - pz * pz);
    }
    E[0][0] = (1.0f / 3.0f) * (2.0f * e[0] - e[4]); E[0][1] = 0.5f * e[1]; E[0][2] = 0.5f * e[2];
    E[1][0] = E[0][1]; E[1][1] = (1.0f / 3.0f) * (-e[0] + 2.0f * e[4]); E[1][2] = 0.5f * e[3];
    E[2][0] = E[0][2]; E[2][1] = E[1][2]; E[2][2] = (-1.0f / 3.0f) * (e[0] + e[4]);
}

// Lubrication_RFE_kernel (Lubrication.cu:350-608): F += R_FE : E
__global__ void rfe_kernel(const float4 *pos, int N, BoxF b, const int *nneigh, const int *nlist, int maxn, LubTable t,
                           float rl2, EParams ep, float *F)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float4 pi = pos[i];
    float Ei[3][3];
    strain_block(ep, i, Ei);
    float acc[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < nneigh[i]; ++k) {
        int j = nlist[(int64_t)i * maxn + k];
        float4 pj = pos[j];
        float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        box_min_image(b, dx, dy, dz);
        float d2 = dist2_nc(dx, dy, dz);
        if (!(d2 < rl2)) continue;
        float Ej[3][3];
        strain_block(ep, j, Ej);
        float dist = sqrtf(d2);
        float v[6];
        lub_lookup(t, dist, 10, 6, 232, v);
        const float XG11 = v[0], XG12 = v[1], YG11 = v[2], YG12 = v[3], YH11 = v[4], YH12 = v[5];
        const float r[3] = {dx / dist, dy / dist, dz / dist};
        float Ei_r[3], Ej_r[3];
        for (int a = 0; a < 3; ++a) {
            Ei_r[a] = Ei[a][0] * r[0] + Ei[a][1] * r[1] + Ei[a][2] * r[2];
            Ej_r[a] = Ej[a][0] * r[0] + Ej[a][1] * r[1] + Ej[a][2] * r[2];
        }
        float rEi = r[0] * Ei_r[0] + r[1] * Ei_r[1] + r[2] * Ei_r[2];
        float rEj = r[0] * Ej_r[0] + r[1] * Ej_r[1] + r[2] * Ej_r[2];
        float eEi[3] = {r[2] * Ei_r[1] - r[1] * Ei_r[2], -r[2] * Ei_r[0] + r[0] * Ei_r[2], r[1] * Ei_r[0] - r[0] * Ei_r[1]};
        float eEj[3] = {r[2] * Ej_r[1] - r[1] * Ej_r[2], -r[2] * Ej_r[0] + r[0] * Ej_r[2], r[1] * Ej_r[0] - r[0] * Ej_r[1]};
        const float XG21 = -XG12, YG21 = -YG12, YH21 = YH12;
        for (int a = 0; a < 3; ++a) {
            acc[a] += (XG11 - 2.0f * YG11) * (-rEi) * r[a] + 2.0f * YG11 * (-Ei_r[a]) + (XG21 - 2.0f * YG21) * (-rEj) * r[a]
                      + 2.0f * YG21 * (-Ej_r[a]);
            acc[3 + a] += YH11 * (2.0f * eEi[a]) + YH21 * (2.0f * eEj[a]);
        }
    }
    for (int a = 0; a < 6; ++a) F[6 * (int64_t)i + a] += acc[a];
}

// Lubrication_RSU_kernel (Lubrication.cu:650-911): S -= R_SU U
__global__ void rsu_kernel(const float4 *pos, int N, BoxF b, const int *nneigh, const int *nlist, int maxn, LubTable t,
                           float rl2, const float *U, float *S)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float4 pi = pos[i];
    const float *ui = U + 6 * (int64_t)i, *wi = U + 6 * (int64_t)i + 3;
    float si[5] = {0, 0, 0, 0, 0};
    const int ca[5] = {0, 0, 0, 1, 1}, cb[5] = {0, 1, 2, 2, 1};
    for (int k = 0; k < nneigh[i]; ++k) {
        int j = nlist[(int64_t)i * maxn + k];
        float4 pj = pos[j];
        float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        box_min_image(b, dx, dy, dz);
        float d2 = dist2_nc(dx, dy, dz);
        if (!(d2 < rl2)) continue;
        float dist = sqrtf(d2);
        float v[6];
        lub_lookup(t, dist, 10, 6, 232, v);
        const float XG11 = v[0], XG12 = v[1], YG11 = v[2], YG12 = v[3], YH11 = v[4], YH12 = v[5];
        const float r[3] = {dx / dist, dy / dist, dz / dist};
        const float *uj = U + 6 * (int64_t)j, *wj = U + 6 * (int64_t)j + 3;
        float rdui = r[0] * ui[0] + r[1] * ui[1] + r[2] * ui[2];
        float rduj = r[0] * uj[0] + r[1] * uj[1] + r[2] * uj[2];
        float ewi[3] = {r[2] * wi[1] - r[1] * wi[2], -r[2] * wi[0] + r[0] * wi[2], r[1] * wi[0] - r[0] * wi[1]};
        float ewj[3] = {r[2] * wj[1] - r[1] * wj[2], -r[2] * wj[0] + r[0] * wj[2], r[1] * wj[0] - r[0] * wj[1]};
        for (int c = 0; c < 5; ++c) {
            int a = ca[c], bb = cb[c];
            float third = (a == bb) ? (1.0f / 3.0f) : 0.f;
            si[c] += XG11 * (r[a] * r[bb] - third) * rdui + XG12 * (r[a] * r[bb] - third) * rduj
                     + YG11 * (ui[a] * r[bb] + r[a] * ui[bb] - 2.0f * r[a] * r[bb] * rdui)
                     + YG12 * (uj[a] * r[bb] + r[a] * uj[bb] - 2.0f * r[a] * r[bb] * rduj);
            si[c] += YH11 * (r[a] * ewi[bb] + ewi[a] * r[bb]) + YH12 * (r[a] * ewj[bb] + ewj[a] * r[bb]);
        }
    }
    for (int c = 0; c < 5; ++c) S[5 * (int64_t)i + c] -= si[c];
}

// Lubrication_RSE_kernel (Lubrication.cu:953-1211): S += R_SE : E, cutoff r^2 < 16
__global__ void rse_kernel(const float4 *pos, int N, BoxF b, const int *nneigh, const int *nlist, int maxn, LubTable t,
                           EParams ep, float *S)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float4 pi = pos[i];
    float Ei[3][3];
    strain_block(ep, i, Ei);
    float Si[5] = {0, 0, 0, 0, 0};
    const int ca[5] = {0, 0, 0, 1, 1}, cb[5] = {0, 1, 2, 2, 1};
    for (int k = 0; k < nneigh[i]; ++k) {
        int j = nlist[(int64_t)i * maxn + k];
        float4 pj = pos[j];
        float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        box_min_image(b, dx, dy, dz);
        float d2 = dist2_nc(dx, dy, dz);
        if (!((double)d2 < 16.0)) continue;
        float Ej[3][3];
        strain_block(ep, j, Ej);
        float dist = sqrtf(d2);
        float v[6];
        lub_lookup(t, dist, 16, 6, 232, v);
        const float XM11 = v[0], XM12 = v[1], YM11 = v[2], YM12 = v[3], ZM11 = v[4], ZM12 = v[5];
        const float r[3] = {dx / dist, dy / dist, dz / dist};
        float Ei_r[3], Ej_r[3];
        for (int a = 0; a < 3; ++a) {
            Ei_r[a] = Ei[a][0] * r[0] + Ei[a][1] * r[1] + Ei[a][2] * r[2];
            Ej_r[a] = Ej[a][0] * r[0] + Ej[a][1] * r[1] + Ej[a][2] * r[2];
        }
        float rEi = r[0] * Ei_r[0] + r[1] * Ei_r[1] + r[2] * Ei_r[2];
        float rEj = r[0] * Ej_r[0] + r[1] * Ej_r[1] + r[2] * Ej_r[2];
        for (int c = 0; c < 5; ++c) {
            int a = ca[c], bb = cb[c];
            float third = (a == bb) ? (1.0f / 3.0f) : 0.f;
            float del = (a == bb) ? 1.f : 0.f;
            Si[c] += 1.5f * XM11 * (r[a] * r[bb] - third) * rEi + 1.5f * XM12 * (r[a] * r[bb] - third) * rEj
                     + 0.5f * YM11 * (2.0f * r[a] * Ei_r[bb] + 2.0f * r[bb] * Ei_r[a] - 4.0f * rEi * r[a] * r[bb])
                     + 0.5f * YM12 * (2.0f * r[a] * Ej_r[bb] + 2.0f * r[bb] * Ej_r[a] - 4.0f * rEj * r[a] * r[bb])
                     + 0.5f * ZM11 * (2.f * Ei[a][bb] + (del + r[a] * r[bb]) * rEi - 2.0f * r[a] * Ei_r[bb] - 2.0f * r[bb] * Ei_r[a])
                     + 0.5f * ZM12 * (2.f * Ej[a][bb] + (del + r[a] * r[bb]) * rEj - 2.0f * r[a] * Ej_r[bb] - 2.0f * r[bb] * Ej_r[a]);
        }
    }
    for (int c = 0; c < 5; ++c) S[5 * (int64_t)i + c] += Si[c];
}

static EParams make_ep(fsd_ctx *ctx, float shear_rate, const float *ori)
{
    EParams ep;
    ep.shear_rate = shear_rate;
    ep.B2 = (ori != nullptr) ? ctx->p.sqm_B2 : 0.f;
    ep.mask2 = ctx->d_mask2;
    ep.ori = ori;
    return ep;
}

int k_rfe(fsd_ctx *ctx, const float4 *pos, float shear_rate, const float *ori, float *F)
{
    LubTable t{ctx->d_res_dist, ctx->d_res_vals, ctx->table_min, ctx->table_dr};
    hipLaunchKernelGGL(rfe_kernel, dim3(nblk(ctx->N, 128)), dim3(128), 0, ctx->stream, pos, ctx->N, ctx->box,
                       ctx->d_nneigh, ctx->d_nlist, ctx->maxn, t, ctx->p.rlub * ctx->p.rlub,
                       make_ep(ctx, shear_rate, ori), F);
    FSD_HIP(hipGetLastError());
    return 0;
}

int k_rsu(fsd_ctx *ctx, const float4 *pos, const float *U, float *S)
{
    LubTable t{ctx->d_res_dist, ctx->d_res_vals, ctx->table_min, ctx->table_dr};
    hipLaunchKernelGGL(rsu_kernel, dim3(nblk(ctx->N, 128)), dim3(128), 0, ctx->stream, pos, ctx->N, ctx->box,
                       ctx->d_nneigh, ctx->d_nlist, ctx->maxn, t, ctx->p.rlub * ctx->p.rlub, U, S);
    FSD_HIP(hipGetLastError());
    return 0;
}

int k_rse(fsd_ctx *ctx, const float4 *pos, float shear_rate, const float *ori, float *S)
{
    LubTable t{ctx->d_res_dist, ctx->d_res_vals, ctx->table_min, ctx->table_dr};
    hipLaunchKernelGGL(rse_kernel, dim3(nblk(ctx->N, 128)), dim3(128), 0, ctx->stream, pos, ctx->N, ctx->box,
                       ctx->d_nneigh, ctx->d_nlist, ctx->maxn, t, make_ep(ctx, shear_rate, ori), S);
    FSD_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------ force law
// Stokes_SetForce_manually_kernel (Helper_Stokes.cu:65-172)
__global__ void force_law_kernel(const float4 *pos, int N, BoxF b, const int *nneigh, const int *nlist, int maxn,
                                 float ndsr, float k_n, float kappa, float beta, float epsq, float T_ext, float *F)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float4 pi = pos[i];
    const float h_rough = sqrtf(epsq), rcol = 2.0f + 1.0f * h_rough, F0 = 1.0f / ndsr, Ham = F0 * beta;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    for (int k = 0; k < nneigh[i]; ++k) {
        int j = nlist[(int64_t)i * maxn + k];
        float4 pj = pos[j];
        float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
        box_min_image(b, dx, dy, dz);
        float dist = sqrtf(dist2_nc(dx, dy, dz));
        float gap = dist - rcol;
        float mag = 0.f;
        if (gap >= 0.f && dist <= 2.0f + 10.0f / kappa) mag = Ham / (12.f * (gap * gap + epsq)) - F0 * expf(-kappa * gap);
        if (gap < 0.f) mag = Ham / (12.f * epsq) - F0 - k_n * fabsf(gap);
        fx += mag * (dx / dist); fy += mag * (dy / dist); fz += mag * (dz / dist);
    }
    F[6 * (int64_t)i] = fx; F[6 * (int64_t)i + 1] = fy; F[6 * (int64_t)i + 2] = fz;
    F[6 * (int64_t)i + 3] = 0.f; F[6 * (int64_t)i + 4] = 0.f; F[6 * (int64_t)i + 5] = T_ext;
}

int k_force_law(fsd_ctx *ctx, const float4 *pos, float T_ext, float *F)
{
    const fsd_params &p = ctx->p;
    hipLaunchKernelGGL(force_law_kernel, dim3(nblk(ctx->N, 128)), dim3(128), 0, ctx->stream, pos, ctx->N, ctx->box,
                       ctx->d_nneigh, ctx->d_nlist, ctx->maxn, p.ndsr, p.k_n, p.kappa, p.beta_AF, p.epsq, T_ext, F);
    FSD_HIP(hipGetLastError());
    return 0;
}

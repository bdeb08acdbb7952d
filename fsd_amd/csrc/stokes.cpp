// Host side of the MI355X FSD step: context set-up (Stokes::setParams,
// setResistanceTable, setSparseMath, AllocateWorkSpaces), the preconditioner
// build around the host RCM, the GMRES driver (FP32 Arnoldi on the device,
// FP64 Hessenberg/Givens on the host) and the step sequencer
// (Stokes::integrateStepOne -> Stokes_StepOne), exported through fsd.h.
#include "fsd_internal.h"
#include "sweep_plan.h"
#include "lanczos_host.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

extern "C" {
int fsd_ewald_table_size(float ewald_cut, float ewald_dr);
void fsd_ewald_table_fill(float xi, int nR, float *tab);
void fsd_ewald_self(float xi, float self_out[2]);
}

// ------------------------------------------------------------ profiling
static const char *kProfNames[PROF_COUNT] = {"spread", "green", "contract", "fft", "realspace", "rfu",
                                             "trsv", "ichol", "gmres_vec", "nlist", "precond_build", "sweep",
                                             "plan_host", "finish_host", "ljit", "learned_net", "jacobi"};

void prof_host(fsd_ctx *ctx, int slot, double ms)
{
    if (!(ctx->prof & (1u << slot))) return;
    ctx->slots[slot].ms += ms;
    ctx->slots[slot].launches += 1;
}

void prof_begin(fsd_ctx *ctx, int slot, hipEvent_t *ev)
{
    *ev = nullptr;
    if (!(ctx->prof & (1u << slot))) return;
    if (ctx->slots[slot].seen++ % ctx->prof_every != 0) return;
    hipEventCreate(ev);
    hipEventRecord(*ev, ctx->stream);
}

void prof_end_n(fsd_ctx *ctx, int slot, hipEvent_t ev, int n)
{
    if (!(ctx->prof & (1u << slot)) || !ev) return;
    hipEvent_t e2;
    hipEventCreate(&e2);
    hipEventRecord(e2, ctx->stream);
    ctx->slots[slot].pending.push_back({ev, e2, n});
}

void prof_end(fsd_ctx *ctx, int slot, hipEvent_t ev) { prof_end_n(ctx, slot, ev, 1); }

static void prof_collect(fsd_ctx *ctx)
{
    hipStreamSynchronize(ctx->stream);
    for (auto &s : ctx->slots) {
        for (auto &pr : s.pending) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, pr.b, pr.e);
            s.ms += ms;
            s.launches += pr.n;
            hipEventDestroy(pr.b);
            hipEventDestroy(pr.e);
        }
        s.pending.clear();
    }
}

// ------------------------------------------------------------ parameters
// Stokes::setParams (Stokes.cc:131-261) and the spreading constants of
// gpu_stokes_Mwave_wrap (Mobility.cu:930-935).
static int pse_params(fsd_ctx *ctx)
{
    const fsd_params &p = ctx->p;
    fsd_pse_info &o = ctx->pse;
    memset(&o, 0, sizeof(o));
    const float xi = p.xi, error = p.error;
    o.rcut = sqrtf(-logf(error)) / xi;
    const int kmax = (int)(2.0 * sqrtf(-logf(error)) * xi) + 1;
    if (o.rcut > p.Lx / 2.0 || o.rcut > p.Ly / 2.0 || o.rcut > p.Lz / 2.0) {
        char buf[256];
        snprintf(buf, sizeof buf, "real space Ewald cutoff %g exceeds half the box; increase xi", o.rcut);
        ctx->err = buf;
        return FSD_E_PARAM;
    }
    const float L[3] = {p.Lx, p.Ly, p.Lz};
    int n[3];
    std::vector<int> good;
    for (int p2 = 1; p2 <= 512; p2 *= 2)
        for (int p3 = 1; p2 * p3 <= 512; p3 *= 3)
            for (int p5 = 1; p2 * p3 * p5 <= 512; p5 *= 5)
                if (p2 * p3 * p5 >= 8) good.push_back(p2 * p3 * p5);
    std::sort(good.begin(), good.end());
    for (int d = 0; d < 3; ++d) {
        int m = (int)(kmax * L[d] / (2.0 * 3.1415926536) * 2.0) + 1;
        auto it = std::lower_bound(good.begin(), good.end(), m);
        if (it == good.end()) { ctx->err = "grid larger than 512 points per axis"; return FSD_E_PARAM; }
        n[d] = *it;
    }
    o.Nx = n[0]; o.Ny = n[1]; o.Nz = n[2];
    const float gamma = p.max_strain, gamma2 = gamma * gamma;
    const float lambda = (float)(1.0 + gamma2 / 2.0 + gamma * sqrtf((float)(1.0 + gamma2 / 4.0)));
    o.gridh_x = p.Lx / (float)o.Nx; o.gridh_y = p.Ly / (float)o.Ny; o.gridh_z = p.Lz / (float)o.Nz;
    float gm = 1.0f;
    while (erfcf(gm / sqrtf((float)(2.0 * lambda))) > error) gm = (float)(gm + 0.01);
    o.gaussm = gm;
    o.P = (int)(gm * gm / 3.1415926536) + 1;
    const float w = (float)(o.P * o.gridh_x / 2.0);
    const float xisq = xi * xi;
    o.eta = (float)((2.0 * w / gm) * (2.0 * w / gm) * xisq);
    if (o.P > std::min(o.Nx, std::min(o.Ny, o.Nz))) { ctx->err = "Gaussian support exceeds the grid"; return FSD_E_PARAM; }
    o.ewald_n = (int)(o.rcut / 0.001f - 1.0f);
    float self[2];
    fsd_ewald_self(xi, self);
    o.self_x = self[0]; o.self_y = self[1];
    o.quadW = o.gridh_x * o.gridh_y * o.gridh_z;
    o.prefac = (float)((2.0 * xisq / 3.1415926536 / o.eta) * sqrtf((float)(2.0 * xisq / 3.1415926536 / o.eta)));
    o.expfac = (float)(2.0 * xisq / o.eta);
    return 0;
}

static void set_box(fsd_ctx *ctx, float Lx, float Ly, float Lz, float xy, float xz, float yz)
{
    BoxF &b = ctx->box;
    b.Lx = Lx; b.Ly = Ly; b.Lz = Lz; b.xy = xy; b.xz = xz; b.yz = yz;
    b.iLx = 1.0f / Lx; b.iLy = 1.0f / Ly; b.iLz = 1.0f / Lz;
}

// cells along each axis for the neighbour list (nearest plane distances)
static void cell_dims(fsd_ctx *ctx)
{
    const BoxF &b = ctx->box;
    const double a1[3] = {b.Lx, 0, 0}, a2[3] = {b.xy * b.Ly, b.Ly, 0}, a3[3] = {b.xz * b.Lz, b.yz * b.Lz, b.Lz};
    auto cross = [](const double *u, const double *v, double *o) {
        o[0] = u[1] * v[2] - u[2] * v[1]; o[1] = u[2] * v[0] - u[0] * v[2]; o[2] = u[0] * v[1] - u[1] * v[0];
    };
    double c23[3], c31[3], c12[3];
    cross(a2, a3, c23); cross(a3, a1, c31); cross(a1, a2, c12);
    const double V = (double)b.Lx * b.Ly * b.Lz;
    auto nrm = [](const double *u) { return std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]); };
    const double d[3] = {V / nrm(c23), V / nrm(c31), V / nrm(c12)};
    const double rl = ctx->pse.rcut + ctx->p.r_buff;
    ctx->use_cells = true;
    for (int a = 0; a < 3; ++a) {
        ctx->ncell[a] = (int)std::floor(d[a] / rl);
        if (ctx->ncell[a] < 3) ctx->use_cells = false;
    }
    ctx->ncells = ctx->use_cells ? ctx->ncell[0] * ctx->ncell[1] * ctx->ncell[2] : 0;
}

#define CK(call)                                                                                    \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) {                                                                     \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);                           \
            return FSD_E_HIP;                                                               \
        }                                                                                           \
    } while (0)

#define RC(call)                                                                                    \
    do {                                                                                            \
        int r_ = (call);                                                                            \
        if (r_ < 0) return r_;                                                                      \
    } while (0)

static int alloc_all(fsd_ctx *ctx)
{
    const int N = ctx->N;
    const int64_t n17 = 17 * (int64_t)N;
    // neighbour list capacity from the density (list radius rcut + r_buff)
    const double rl = ctx->pse.rcut + ctx->p.r_buff;
    const double rho = N / ((double)ctx->box.Lx * ctx->box.Ly * ctx->box.Lz);
    int est = (int)(4.18879 * rl * rl * rl * rho * 1.6) + 32;
    est = std::min(est, N);
    ctx->maxn = ((est + 31) / 32) * 32;
    CK(hipMalloc(&ctx->d_nneigh, sizeof(int) * N));
    CK(hipMalloc(&ctx->d_nlist, sizeof(int) * (size_t)N * ctx->maxn));
    if (ctx->use_cells) {
        CK(hipMalloc(&ctx->d_cell_of, sizeof(int) * N));
        CK(hipMalloc(&ctx->d_cell_count, sizeof(int) * ctx->ncells));
        CK(hipMalloc(&ctx->d_cell_start, sizeof(int) * (ctx->ncells + 1)));
        CK(hipMalloc(&ctx->d_cell_list, sizeof(int) * N));
    }
    CK(hipMalloc(&ctx->d_flags, sizeof(int) * 8));
    CK(hipMalloc(&ctx->d_pos_last, sizeof(float4) * (size_t)N));
    CK(hipHostMalloc((void **)&ctx->h_flags, sizeof(int) * 8, hipHostMallocDefault));
    // lubrication records
    ctx->maxlub = 48;
    CK(hipMalloc(&ctx->d_lub_n, sizeof(int) * N));
    CK(hipMalloc(&ctx->d_lub_j, sizeof(int) * (size_t)N * ctx->maxlub));
    CK(hipMalloc(&ctx->d_lub_a, sizeof(float4) * (size_t)N * ctx->maxlub));
    CK(hipMalloc(&ctx->d_lub_b, sizeof(float4) * (size_t)N * ctx->maxlub));
    CK(hipMalloc(&ctx->d_lub_self, sizeof(float) * 36 * (size_t)N));
    // grids and FFT
    const fsd_pse_info &ps = ctx->pse;
    ctx->M = ps.Nx * ps.Ny * ps.Nz;
    ctx->Mh = ps.Nx * ps.Ny * (ps.Nz / 2 + 1);
    CK(hipMalloc(&ctx->d_grid, sizeof(float) * 11 * (size_t)ctx->M));
    CK(hipMalloc(&ctx->d_spec, sizeof(float2) * 11 * (size_t)ctx->Mh));
    CK(hipMalloc(&ctx->d_gridk, sizeof(float4) * (size_t)ctx->Mh));
    ctx->tile = 8;
    ctx->tiles[0] = (ps.Nx + 7) / 8; ctx->tiles[1] = (ps.Ny + 7) / 8; ctx->tiles[2] = (ps.Nz + 7) / 8;
    ctx->ntiles = ctx->tiles[0] * ctx->tiles[1] * ctx->tiles[2];
    CK(hipMalloc(&ctx->d_porig, sizeof(int) * 3 * N));
    CK(hipMalloc(&ctx->d_ocell, sizeof(int) * N));
    CK(hipMalloc(&ctx->d_olist, sizeof(int) * N));
    CK(hipMalloc(&ctx->d_ocount, sizeof(int) * (size_t)ctx->M));
    CK(hipMalloc(&ctx->d_ostart, sizeof(int) * ((size_t)ctx->M + 1)));
    CK(hipMalloc(&ctx->d_scan_tmp, sizeof(int) * 2 * (scan_excl_chunks(ctx->M) + 1)));
    if (ctx->tile + ps.P - 1 > 32) { ctx->err = "spreading support P too large for the tile halo"; return FSD_E_PARAM; }
    ctx->contract_lds = k_contract_lds(ctx);
    // FSD_TRSV_LEVELS=1 runs the triangular sweeps one launch per level; IC(0) applied by Jacobi
    // iterations needs no sweep plans (the learned mode keeps them for an exact IC(0) fallback)
    {
        const char *lv = getenv("FSD_TRSV_LEVELS");
        if ((lv && lv[0] == '1') || ctx->ic_jac_k > 0) ctx->tp_R = 0;
        else RC(k_sweep_setup(ctx));
        // FSD_SWEEP_JOINT=1 runs both right-hand sides of a sweep in one workgroup
        const char *jt = getenv("FSD_SWEEP_JOINT");
        ctx->tp_joint = jt && jt[0] == '1';
    }
    CK(hipMalloc(&ctx->d_wave, sizeof(float) * 11 * (size_t)N));
    CK(hipMalloc(&ctx->d_rec, sizeof(float4) * 3 * (size_t)N));
    {
        size_t lengths[3] = {(size_t)ps.Nz, (size_t)ps.Ny, (size_t)ps.Nx};
        if (rocfft_plan_create(&ctx->plan_fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                               rocfft_precision_single, 3, lengths, 11, nullptr) != rocfft_status_success ||
            rocfft_plan_create(&ctx->plan_inv, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                               rocfft_precision_single, 3, lengths, 11, nullptr) != rocfft_status_success) {
            ctx->err = "rocfft_plan_create failed";
            return FSD_E_HIP;
        }
        size_t w1 = 0, w2 = 0;
        rocfft_plan_get_work_buffer_size(ctx->plan_fwd, &w1);
        rocfft_plan_get_work_buffer_size(ctx->plan_inv, &w2);
        ctx->fftwork_size = std::max(w1, w2);
        // the mobility's transforms: (y, z) by rocFFT, batched over the 11 fields and the x planes (the
        // layout of the 3-D half spectrum), x and the Green's function in one kernel when Nx = 8 B,
        // 5 <= B <= 11.  FSD_FFT_3D=1 keeps the 3-D plans and the separate Green's function.
        const char *f3 = getenv("FSD_FFT_3D");
        const int B = ps.Nx / 8;
        if (!(f3 && f3[0] == '1') && ps.Nx % 8 == 0 && B >= 5 && B <= 11) {
            size_t len2[2] = {(size_t)ps.Nz, (size_t)ps.Ny};
            const size_t batch = 11 * (size_t)ps.Nx;
            if (rocfft_plan_create(&ctx->plan_fwd2, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                                   rocfft_precision_single, 2, len2, batch, nullptr) != rocfft_status_success ||
                rocfft_plan_create(&ctx->plan_inv2, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                                   rocfft_precision_single, 2, len2, batch, nullptr) != rocfft_status_success) {
                ctx->err = "rocfft_plan_create (2-D) failed";
                return FSD_E_HIP;
            }
            rocfft_plan_get_work_buffer_size(ctx->plan_fwd2, &w1);
            rocfft_plan_get_work_buffer_size(ctx->plan_inv2, &w2);
            ctx->fftwork_size = std::max(ctx->fftwork_size, std::max(w1, w2));
            ctx->fx_B = B;
        }
        if (ctx->fftwork_size) CK(hipMalloc(&ctx->d_fftwork, ctx->fftwork_size));
        rocfft_execution_info_create(&ctx->info_fwd);
        rocfft_execution_info_create(&ctx->info_inv);
        rocfft_execution_info_set_stream(ctx->info_fwd, ctx->stream);
        rocfft_execution_info_set_stream(ctx->info_inv, ctx->stream);
        if (ctx->fftwork_size) {
            rocfft_execution_info_set_work_buffer(ctx->info_fwd, ctx->d_fftwork, ctx->fftwork_size);
            rocfft_execution_info_set_work_buffer(ctx->info_inv, ctx->d_fftwork, ctx->fftwork_size);
        }
    }
    // Ewald table
    ctx->ewald_nR = ps.ewald_n + 1;
    ctx->h_ewald.resize(8 * (size_t)ctx->ewald_nR);
    fsd_ewald_table_fill(ctx->p.xi, ctx->ewald_nR, ctx->h_ewald.data());
    CK(hipMalloc(&ctx->d_ewald, sizeof(float) * ctx->h_ewald.size()));
    CK(hipMemcpy(ctx->d_ewald, ctx->h_ewald.data(), sizeof(float) * ctx->h_ewald.size(), hipMemcpyHostToDevice));
    CK(hipMalloc(&ctx->d_res_dist, sizeof(float) * 1000));
    CK(hipMalloc(&ctx->d_res_vals, sizeof(float) * 22000));
    // preconditioner
    CK(hipMalloc(&ctx->d_pn, sizeof(int) * N));
    CK(hipMalloc(&ctx->d_plist, sizeof(int) * FSD_PSTRIDE * (size_t)N));
    CK(hipMalloc(&ctx->d_pless, sizeof(int) * N));
    CK(hipMalloc(&ctx->d_perm, sizeof(int) * N));
    CK(hipMalloc(&ctx->d_iperm, sizeof(int) * N));
    CK(hipMalloc(&ctx->d_lrow, sizeof(int) * (N + 1)));
    CK(hipMalloc(&ctx->d_urow, sizeof(int) * (N + 1)));
    ctx->nlow_cap = 4 * (int64_t)N + 64;
    CK(hipMalloc(&ctx->d_lcol, sizeof(int) * ctx->nlow_cap));
    CK(hipMalloc(&ctx->d_ucol, sizeof(int) * ctx->nlow_cap));
    CK(hipMalloc(&ctx->d_upos, sizeof(int) * ctx->nlow_cap));
    CK(hipMalloc(&ctx->d_Ablk, sizeof(float) * 36 * ctx->nlow_cap));
    CK(hipMalloc(&ctx->d_Lblk, sizeof(float) * 36 * ctx->nlow_cap));
    if (ctx->precond_mode != 0) {
        CK(hipMalloc(&ctx->d_Fblk, sizeof(float) * 36 * ctx->nlow_cap));
        CK(hipMalloc(&ctx->d_lrowidx, sizeof(int) * ctx->nlow_cap));
        CK(hipMalloc(&ctx->d_jac, sizeof(float) * 24 * (size_t)N));
    }
    if (ctx->p.factor_sweeps > 0) CK(hipMalloc(&ctx->d_Lblk2, sizeof(float) * 36 * ctx->nlow_cap));
    if (ctx->p.factor_sweeps > 0) CK(hipMalloc(&ctx->d_Ldiag2, sizeof(float) * 36 * (size_t)N));
    CK(hipMalloc(&ctx->d_Adiag, sizeof(float) * 36 * (size_t)N));
    CK(hipMalloc(&ctx->d_Ldiag, sizeof(float) * 36 * (size_t)N));
    CK(hipMalloc(&ctx->d_lev_f, sizeof(int) * N));
    CK(hipMalloc(&ctx->d_lev_b, sizeof(int) * N));
    CK(hipMalloc(&ctx->d_tp_Z, sizeof(float) * 36 * (size_t)N));
    CK(hipMalloc(&ctx->d_tp_fseq, sizeof(int) * (size_t)N));
    CK(hipMalloc(&ctx->d_tp_bseq, sizeof(int) * (size_t)N));
    CK(hipMalloc(&ctx->d_tp_fplace, sizeof(int) * 3 * (size_t)N));   // a step holds one row at least
    CK(hipMalloc(&ctx->d_tp_bplace, sizeof(int) * 3 * (size_t)N));
    for (float4 **pp : {&ctx->d_tp_finp, &ctx->d_tp_binp, &ctx->d_tp_fres, &ctx->d_tp_bres})
        CK(hipMalloc(pp, sizeof(float4) * 3 * (size_t)N));
    CK(hipMalloc(&ctx->d_tmp6a, sizeof(float) * 12 * (size_t)N));
    CK(hipMalloc(&ctx->d_tmp6b, sizeof(float) * 12 * (size_t)N));
    CK(hipMalloc(&ctx->d_g12, sizeof(float) * 12 * (size_t)N));
    // GMRES
    const int R = ctx->p.gmres_restart;
    CK(hipMalloc(&ctx->d_V, sizeof(float) * n17 * (R + 1)));
    CK(hipMalloc(&ctx->d_w, sizeof(float) * n17));
    CK(hipMalloc(&ctx->d_t, sizeof(float) * n17));
    CK(hipMalloc(&ctx->d_t2, sizeof(float) * n17));
    // dot partials of up to kmax vectors, then 2 kmax + 2 doubles of results (two projection passes, a norm)
    ctx->part_blocks = (int)((n17 + 4095) / 4096);
    ctx->kmax = std::max(64, ctx->p.gmres_restart + 1);
    CK(hipMalloc(&ctx->d_part, sizeof(double) * ((size_t)ctx->part_blocks * ctx->kmax + 2 * ctx->kmax + 2)));
    // pinned: two slots of 2 kmax + 2 doubles (Hessenberg columns read back while the next
    // iteration runs), then kmax solution coefficients
    CK(hipHostMalloc((void **)&ctx->h_part, sizeof(double) * (2 * (2 * ctx->kmax + 2) + ctx->kmax),
                     hipHostMallocDefault));
    for (hipEvent_t *e : {&ctx->gm_ev[0], &ctx->gm_ev[1]}) CK(hipEventCreateWithFlags(e, hipEventDisableTiming));
    CK(hipMalloc(&ctx->d_rhs, sizeof(float) * n17));
    CK(hipMalloc(&ctx->d_sol, sizeof(float) * n17));
    CK(hipMemset(ctx->d_sol, 0, sizeof(float) * n17));
    if (ctx->p.guess >= 1) {
        CK(hipMalloc(&ctx->d_sol_prev, sizeof(float) * n17));
        CK(hipMemset(ctx->d_sol_prev, 0, sizeof(float) * n17));
    }
    if (ctx->p.guess >= 2) {
        CK(hipMalloc(&ctx->d_sol_prev2, sizeof(float) * n17));
        CK(hipMemset(ctx->d_sol_prev2, 0, sizeof(float) * n17));
    }
    CK(hipMalloc(&ctx->d_force, sizeof(float) * 6 * (size_t)N));
    // squirmer masks (Stokes.cc:1318-1346, index == tag)
    std::vector<float> m1(N), m2(N);
    for (int i = 0; i < N; ++i) {
        const bool first = (uint32_t)i < ctx->p.N_mix;
        m1[i] = first ? 1.0f : ctx->p.coef_B1_mask;
        m2[i] = first ? 1.0f : ctx->p.coef_B2_mask;
    }
    CK(hipMalloc(&ctx->d_mask1, sizeof(float) * N));
    CK(hipMalloc(&ctx->d_mask2, sizeof(float) * N));
    CK(hipMemcpy(ctx->d_mask1, m1.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    CK(hipMemcpy(ctx->d_mask2, m2.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    return 0;
}

static void free_all(fsd_ctx *ctx)
{
    lp_free(ctx);
    void *ptrs[] = {ctx->d_nneigh, ctx->d_nlist, ctx->d_cell_of, ctx->d_cell_count, ctx->d_cell_start,
                    ctx->d_cell_list, ctx->d_flags, ctx->d_pos_last, ctx->d_lub_n, ctx->d_lub_j, ctx->d_lub_a, ctx->d_lub_b,
                    ctx->d_lub_self, ctx->d_grid, ctx->d_spec, ctx->d_gridk, ctx->d_fftwork, ctx->d_porig,
                    ctx->d_ocell, ctx->d_olist, ctx->d_ocount, ctx->d_ostart, ctx->d_scan_tmp, ctx->d_wave, ctx->d_rec,
                    ctx->d_ewald, ctx->d_res_dist, ctx->d_res_vals, ctx->d_pn, ctx->d_plist, ctx->d_pless,
                    ctx->d_perm, ctx->d_iperm, ctx->d_lrow, ctx->d_urow, ctx->d_lcol, ctx->d_ucol, ctx->d_upos,
                    ctx->d_Ablk, ctx->d_Lblk, ctx->d_Adiag, ctx->d_Ldiag, ctx->d_lev_f, ctx->d_lev_b, ctx->d_tmp6a,
                    ctx->d_tmp6b, ctx->d_g12, ctx->d_V, ctx->d_w, ctx->d_t, ctx->d_t2, ctx->d_part, ctx->d_rhs, ctx->d_sol, ctx->d_sol_prev, ctx->d_sol_prev2,
                    ctx->d_force, ctx->d_mask1, ctx->d_mask2, ctx->d_tp_finfo, ctx->d_tp_binfo,
                    ctx->d_tp_fbsrc, ctx->d_tp_bbsrc, ctx->d_tp_fblk, ctx->d_tp_bblk,
                    ctx->d_tp_finp, ctx->d_tp_binp, ctx->d_tp_fres, ctx->d_tp_bres, ctx->d_tp_fseq, ctx->d_tp_bseq, ctx->d_tp_fplace, ctx->d_tp_bplace,
                    ctx->d_tp_Z, ctx->d_Fblk, ctx->d_lrowidx, ctx->d_jac, ctx->d_Lblk2, ctx->d_Ldiag2,
                    ctx->d_x64, ctx->d_r64, ctx->d_ybot64, ctx->d_norm64, ctx->d_refx32, ctx->d_refy32,
                    ctx->d_refr32, ctx->d_refd,
                    ctx->d_Z, ctx->d_lanV, ctx->d_lanw, ctx->d_lany0, ctx->d_lan_h, ctx->d_noise, ctx->d_posp,
                    ctx->d_rfd_rhs, ctx->d_rfd_sol};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (ctx->h_flags) hipHostFree(ctx->h_flags);
    if (ctx->h_part) hipHostFree(ctx->h_part);
    for (hipEvent_t e : ctx->gm_ev)
        if (e) hipEventDestroy(e);
    if (ctx->h_lan) hipHostFree(ctx->h_lan);
    if (ctx->plan_fwd) rocfft_plan_destroy(ctx->plan_fwd);
    if (ctx->plan_inv) rocfft_plan_destroy(ctx->plan_inv);
    if (ctx->plan_fwd2) rocfft_plan_destroy(ctx->plan_fwd2);
    if (ctx->plan_inv2) rocfft_plan_destroy(ctx->plan_inv2);
    if (ctx->info_fwd) rocfft_execution_info_destroy(ctx->info_fwd);
    if (ctx->info_inv) rocfft_execution_info_destroy(ctx->info_inv);
    for (auto &s : ctx->slots)
        for (auto &pr : s.pending) { hipEventDestroy(pr.b); hipEventDestroy(pr.e); }
}

// ------------------------------------------------------------ preconditioner
static int grow_lower(fsd_ctx *ctx, int64_t need)
{
    if (need <= ctx->nlow_cap) return 0;
    const int64_t cap = need + need / 4 + 64;
    CK(hipStreamSynchronize(ctx->stream));
    for (int *q : {ctx->d_lcol, ctx->d_ucol, ctx->d_upos}) hipFree(q);
    for (float *q : {ctx->d_Ablk, ctx->d_Lblk}) hipFree(q);
    for (int **pp : {&ctx->d_lcol, &ctx->d_ucol, &ctx->d_upos}) CK(hipMalloc(pp, sizeof(int) * cap));
    for (float **pp : {&ctx->d_Ablk, &ctx->d_Lblk}) CK(hipMalloc(pp, sizeof(float) * 36 * cap));
    if (ctx->d_Fblk) {
        hipFree(ctx->d_Fblk);
        hipFree(ctx->d_lrowidx);
        CK(hipMalloc(&ctx->d_Fblk, sizeof(float) * 36 * cap));
        CK(hipMalloc(&ctx->d_lrowidx, sizeof(int) * cap));
    }
    if (ctx->d_Lblk2) {
        hipFree(ctx->d_Lblk2);
        CK(hipMalloc(&ctx->d_Lblk2, sizeof(float) * 36 * cap));
    }
    ctx->nlow_cap = cap;
    return 0;
}

// device operands of the sweep task plans (lanes grow with the number of steps)
static int grow_tasks(fsd_ctx *ctx)
{
    const int64_t lf = (int64_t)ctx->pch.tp_f.nsteps * TP_LANES, lb = (int64_t)ctx->pch.tp_b.nsteps * TP_LANES;
    auto grow = [ctx](int64_t need, int64_t &cap, int2 *&info, int2 *&bsrc, float4 *&blk) -> int {
        if (need <= cap) return 0;
        const int64_t c = need + need / 4 + TP_LANES;
        CK(hipStreamSynchronize(ctx->stream));
        for (void *q : {(void *)info, (void *)bsrc, (void *)blk})
            if (q) hipFree(q);
        CK(hipMalloc(&info, sizeof(int2) * c));
        CK(hipMalloc(&bsrc, sizeof(int2) * c));
        CK(hipMalloc(&blk, sizeof(float4) * 9 * c));
        cap = c;
        return 0;
    };
    RC(grow(lf, ctx->tp_cap_f, ctx->d_tp_finfo, ctx->d_tp_fbsrc, ctx->d_tp_fblk));
    RC(grow(lb, ctx->tp_cap_b, ctx->d_tp_binfo, ctx->d_tp_bbsrc, ctx->d_tp_bblk));
    return 0;
}

// Precondition_Wrap (Precondition.cu:1326-1497) in three phases: fetch (the
// pruned lists on the GPU, copied to the host), plan (host only: RCM, block
// adjacency, level schedules, sweep plans; it touches no HIP state, so a step
// runs it on a second thread while independent kernels are enqueued) and
// finish (uploads, block values, IC(0), packed sweep operands).
static int precond_fetch(fsd_ctx *ctx, const float4 *pos)
{
    const int N = ctx->N;
    if (!ctx->nlist_valid) { ctx->err = "neighbour list not built"; return FSD_E_STATE; }
    if (!ctx->have_table) { ctx->err = "resistance table not set"; return FSD_E_STATE; }
    prof_begin(ctx, PROF_PRECOND_BUILD, &ctx->pc_ev);
    RC(k_prune(ctx, pos));
    ctx->h_pn.resize(N);
    ctx->h_plstride.resize((size_t)N * FSD_PSTRIDE);
    CK(hipMemcpyAsync(ctx->h_pn.data(), ctx->d_pn, sizeof(int) * N, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipMemcpyAsync(ctx->h_plstride.data(), ctx->d_plist, sizeof(int) * ctx->h_plstride.size(),
                      hipMemcpyDeviceToHost, ctx->stream));
    CK(hipMemcpyAsync(ctx->h_flags + 2, ctx->d_flags + 2, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    if (ctx->h_flags[2]) { ctx->err = "more than 15 neighbours inside the preconditioner cutoff"; return FSD_E_CAPACITY; }
    return 0;
}

static double wall_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void precond_plan(fsd_ctx *ctx, bool keep_order = false)
{
    const double t0 = wall_ms();
    plan_precond(ctx->N, ctx->h_pn.data(), ctx->h_plstride.data(), FSD_PSTRIDE, ctx->tp_R, 4, ctx->pch, keep_order);
    if (!keep_order) ctx->order_nlist = ctx->nlist_builds;   // the list this ordering was computed under
    prof_host(ctx, PROF_PLAN_HOST, wall_ms() - t0);
}

// uploads of the host plan and the pair blocks A (everything a factor needs)
static int precond_upload_blocks(fsd_ctx *ctx, const float4 *pos)
{
    const int N = ctx->N;
    PrecondHost &h = ctx->pch;
    auto up = [ctx](void *d, const void *src, size_t bytes) {
        return hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    };
    RC(grow_lower(ctx, ctx->pch.nlow + 1));
    if (ctx->pch.tp_ok) {
        RC(grow_tasks(ctx));
        // row placements (the lane arrays are built from them on the device)
        const TaskPlan &f = ctx->pch.tp_f, &b = ctx->pch.tp_b;
        CK(up(ctx->d_tp_fseq, f.seq.data(), sizeof(int) * (size_t)N));
        CK(up(ctx->d_tp_bseq, b.seq.data(), sizeof(int) * (size_t)N));
        CK(up(ctx->d_tp_fplace, f.lane.data(), sizeof(int) * (size_t)N));
        CK(up(ctx->d_tp_fplace + N, f.step.data(), sizeof(int) * (size_t)N));
        CK(up(ctx->d_tp_fplace + 2 * (size_t)N, f.send.data(), sizeof(int) * f.send.size()));
        CK(up(ctx->d_tp_bplace, b.lane.data(), sizeof(int) * (size_t)N));
        CK(up(ctx->d_tp_bplace + N, b.step.data(), sizeof(int) * (size_t)N));
        CK(up(ctx->d_tp_bplace + 2 * (size_t)N, b.send.data(), sizeof(int) * b.send.size()));
    }
    CK(up(ctx->d_lev_f, h.itf.data(), sizeof(int) * N));
    CK(up(ctx->d_lev_b, h.itb.data(), sizeof(int) * N));
    CK(up(ctx->d_perm, ctx->pch.perm.data(), sizeof(int) * N));
    CK(up(ctx->d_iperm, h.iperm.data(), sizeof(int) * N));
    CK(up(ctx->d_lrow, h.lrow.data(), sizeof(int) * (N + 1)));
    CK(up(ctx->d_urow, h.urow.data(), sizeof(int) * (N + 1)));
    if (ctx->pch.nlow) {
        CK(up(ctx->d_lcol, h.lcol.data(), sizeof(int) * ctx->pch.nlow));
        CK(up(ctx->d_ucol, h.ucol.data(), sizeof(int) * ctx->pch.nlow));
        CK(up(ctx->d_upos, h.upos.data(), sizeof(int) * ctx->pch.nlow));
    }
    RC(k_build_blocks(ctx, pos));
    return 0;
}

// IC(0) with the relaxer doubling on a non-positive pivot (Precondition.cu:1274-1288, 1455-1489)
// and the operands of its application
static int precond_ichol(fsd_ctx *ctx)
{
    ctx->ichol_attempts = 0;
    float lambda = ctx->ichol_relaxer;
    for (;;) {
        int failed = 0;
        if (ctx->p.factor_sweeps > 0) RC(k_ichol_sweeps(ctx, lambda, ctx->p.factor_sweeps, &failed));
        else RC(k_ichol(ctx, lambda, &failed));
        ctx->ichol_attempts++;
        if (!failed) break;
        lambda *= 2.0f;
        if (lambda > 1.0e7f) { ctx->err = "IC(0) did not succeed for any relaxer"; return FSD_E_ICHOL; }
    }
    ctx->ichol_relaxer = lambda;
    lp_set_active(ctx, false);
    if (ctx->ic_jac_k > 0) RC(k_jacobi_pack(ctx));
    else if (ctx->pch.tp_ok) RC(k_sweep_pack(ctx));
    return 0;
}

static int precond_finish(fsd_ctx *ctx, const float4 *pos)
{
    const double t0 = wall_ms();
    RC(precond_upload_blocks(ctx, pos));
    ctx->lp_step_fallback = 0;
    // learned factor (mode 2), unless this step calibrates IC(0), skips the learned factor after
    // repeated iteration fallbacks, or the factor has a non-positive pivot
    bool learned = false;
    if (ctx->precond_mode == 2) {
        learned = lp_have_weights(ctx) && !ctx->lp_force_ichol;
        if (learned && ctx->lp_skip) {
            ctx->lp_step_fallback = 3;
            learned = false;
        }
        if (learned) {
            int failed = 0;
            RC(lp_forward(ctx, pos, false, ctx->lp_inject, &failed));
            ctx->lp_inject = -1;
            if (failed) {
                ctx->lp_pivot_fallbacks++;
                ctx->lp_step_fallback = 1;
                learned = false;
            }
        }
        lp_set_active(ctx, learned);
        if (learned) ctx->ichol_attempts = 0;
    }
    if (!learned) RC(precond_ichol(ctx));
    prof_host(ctx, PROF_FINISH_HOST, wall_ms() - t0);
    prof_end(ctx, PROF_PRECOND_BUILD, ctx->pc_ev);
    ctx->precond_valid = true;
    return 0;
}

static int build_precond(fsd_ctx *ctx, const float4 *pos)
{
    RC(precond_fetch(ctx, pos));
    precond_plan(ctx);
    return precond_finish(ctx, pos);
}

// ------------------------------------------------------------ GMRES
// Solvers_Saddle with CUSP's gmres semantics (Solvers.cu:55-126): left
// preconditioned restarted GMRES, converged when the preconditioned residual
// estimate |s_{i+1}| <= tol ||b||.  Arnoldi vectors in FP32 on the device,
// classical Gram-Schmidt with one reorthogonalisation and FP64 dot products;
// Hessenberg, Givens rotations and the small triangular solve in FP64 on the host.
static int true_residual(fsd_ctx *ctx, const float4 *pos, const float *b, const float *x, float *out);

static int gmres_core(fsd_ctx *ctx, const float4 *pos, const float *b, float *x, float tol, int maxit, int R,
                      bool right, bool r0_in_t, fsd_solve_stats *st);

// Default: CUSP semantics (left preconditioning, converged when the
// preconditioned residual estimate |s| <= tol ||b||, Solvers.cu:55-126).
// With gmres_true_tol > 0 the solve is right-preconditioned instead, so its
// residual estimate is the true residual ||b - A x||: it stops at
// gmres_true_tol, and the true residual is then recomputed and the solve
// continued (warm) if FP32 drift left it above the tolerance.
// (refined: r0_in_t means refine_first_residual has formed the first FP64 residual)
static int gmres_refined(fsd_ctx *ctx, const float4 *pos, const float *b, float *x, int maxit, int R,
                         fsd_solve_stats *st, bool r0_ready);

static int gmres(fsd_ctx *ctx, const float4 *pos, const float *b, float *x, float tol, int maxit, int R,
                 fsd_solve_stats *st, bool r0_in_t = false)
{
    if (ctx->p.refine && ctx->p.gmres_true_tol > 0.f) return gmres_refined(ctx, pos, b, x, maxit, R, st, r0_in_t);
    const float ttol = ctx->p.gmres_true_tol;
    if (ttol <= 0.f) return gmres_core(ctx, pos, b, x, tol, maxit, R, false, false, st);
    int total = 0, restarts = 0, rc = 0;
    float tres = 0.f, goal = ttol, r0 = -1.f;
    for (int round = 0; round < 6; ++round) {
        rc = gmres_core(ctx, pos, b, x, goal, maxit - total, R, true, r0_in_t && round == 0, st);
        if (rc < 0) return rc;
        if (st && round == 0) r0 = st->initial_residual;
        if (st) { total += st->iterations; restarts += st->restarts; }
        RC(true_residual(ctx, pos, b, x, &tres));
        if (tres <= ttol || total >= maxit) break;
        goal = goal * std::max(0.1f, 0.9f * ttol / tres);
    }
    if (st) {
        st->iterations = total;
        st->restarts = restarts;
        st->true_residual = tres;
        st->initial_residual = r0;
    }
    return (tres <= ttol) ? 0 : FSD_W_NOT_CONVERGED;
}

// FP64-residual iterative refinement (fsd_params.refine): x is accumulated in FP64;
// each round forms r = b - A x with the lubrication product and the subtraction in
// FP64 (k_residual64), stops at |r| <= gmres_true_tol |b|, and otherwise solves
// A d = r with the FP32 right-preconditioned GMRES to a relative tolerance aimed at
// the target (at most 0.5, at least 1e-4, the FP32 floor) and adds d in FP64.
static int refine_alloc(fsd_ctx *ctx)
{
    if (ctx->d_x64) return 0;
    const int64_t n = 17 * (int64_t)ctx->N;
    CK(hipMalloc(&ctx->d_x64, sizeof(double) * n));
    CK(hipMalloc(&ctx->d_r64, sizeof(double) * n));
    CK(hipMalloc(&ctx->d_ybot64, sizeof(double) * 6 * (size_t)ctx->N));
    CK(hipMalloc(&ctx->d_norm64, sizeof(double) * 2));
    for (float **pp : {&ctx->d_refx32, &ctx->d_refy32, &ctx->d_refr32, &ctx->d_refd})
        CK(hipMalloc(pp, sizeof(float) * n));
    return 0;
}

// r = b - A x64 in FP64 (d_r64, and in FP32 d_refr32), nr = |r|^2, |b|^2
static int refine_residual(fsd_ctx *ctx, const float4 *pos, const float *b, double nr[2])
{
    RC(k_residual64(ctx, pos, b, ctx->d_x64, ctx->d_refx32, ctx->d_refy32, ctx->d_ybot64, ctx->d_r64,
                    ctx->d_refr32, ctx->d_norm64));
    CK(hipMemcpyAsync(ctx->h_part, ctx->d_norm64, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    nr[0] = ctx->h_part[0];
    nr[1] = ctx->h_part[1];
    return 0;
}

// the refinement's first residual, formed before the preconditioner exists: x64 = x and r,
// whose norms gmres_refined then takes from ctx->ref_r0 (fsd_step forms them while the host
// plans the preconditioner)
static int refine_first_residual(fsd_ctx *ctx, const float4 *pos, const float *b, const float *x)
{
    RC(refine_alloc(ctx));
    RC(k_to64(ctx, 17 * (int64_t)ctx->N, x, ctx->d_x64));
    return refine_residual(ctx, pos, b, ctx->ref_r0);
}

static int gmres_refined(fsd_ctx *ctx, const float4 *pos, const float *b, float *x, int maxit, int R,
                         fsd_solve_stats *st, bool r0_ready)
{
    const int64_t n = 17 * (int64_t)ctx->N;
    RC(refine_alloc(ctx));
    const double ttol = ctx->p.gmres_true_tol;
    if (!r0_ready) RC(k_to64(ctx, n, x, ctx->d_x64));
    int total = 0, restarts = 0, rounds = 0;
    double rel = 1.0, r0 = -1.0;
    for (;;) {
        double nr[2];
        if (r0_ready) {
            nr[0] = ctx->ref_r0[0];
            nr[1] = ctx->ref_r0[1];
            r0_ready = false;
        } else {
            RC(refine_residual(ctx, pos, b, nr));
        }
        rel = std::sqrt(nr[0]) / std::max(std::sqrt(nr[1]), 1e-300);
        if (r0 < 0.0) r0 = rel;
        if (rel <= ttol || total >= maxit || rounds >= 12) break;
        RC(k_fill(ctx, n, 0.f, ctx->d_refd));
        const float tau = (float)std::min(0.5, std::max(1e-4, 0.5 * ttol / rel));
        fsd_solve_stats si;
        memset(&si, 0, sizeof si);
        const int rc = gmres_core(ctx, pos, ctx->d_refr32, ctx->d_refd, tau, maxit - total, R, true, false, &si);
        if (rc < 0) return rc;
        total += si.iterations;
        restarts += si.restarts;
        RC(k_add64(ctx, n, ctx->d_refd, ctx->d_x64));
        ++rounds;
    }
    RC(k_to32(ctx, n, ctx->d_x64, x));
    if (st) {
        st->iterations = total;
        st->restarts = restarts;
        st->residual = (float)rel;
        st->true_residual = (float)rel;
        st->initial_residual = (float)r0;
    }
    return rel <= ttol ? 0 : FSD_W_NOT_CONVERGED;
}

// Restarted GMRES(R), classical Gram-Schmidt with one reorthogonalisation and
// FP64 dot products; Hessenberg, Givens rotations and the triangular solve in
// FP64 on the host.  left: Krylov space of M^-1 A, residual M^-1 (b - A x);
// right: Krylov space of A M^-1 with Z = M^-1 V kept, residual b - A x.
static int gmres_core(fsd_ctx *ctx, const float4 *pos, const float *b, float *x, float tol, int maxit, int R,
                      bool right, bool r0_in_t, fsd_solve_stats *st)
{
    const int64_t n = 17 * (int64_t)ctx->N;
    float *V = ctx->d_V, *w = ctx->d_w, *t = ctx->d_t;
    if (right && !ctx->d_Z) CK(hipMalloc(&ctx->d_Z, sizeof(float) * n * R));
    float *Z = ctx->d_Z;
    const int KM = ctx->kmax, HS = 2 * KM + 2;
    double *d_h = ctx->d_part + (int64_t)ctx->part_blocks * KM;   // 2 kmax + 2 doubles after the partials
    std::vector<double> H((size_t)(R + 1) * R, 0.0), cs(R), sn(R), g(R + 1), yv(R + 1);
    auto hostsync = [&](int count) -> int {
        CK(hipMemcpyAsync(ctx->h_part, d_h, sizeof(double) * count, hipMemcpyDeviceToHost, ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
        return 0;
    };
    RC(k_nrm2(ctx, n, b, d_h));
    RC(hostsync(1));
    const double bn = std::sqrt(ctx->h_part[0]);
    const double target = (double)tol * bn;
    int iters = 0, restarts = 0;
    double resid = 0.0;
    for (;;) {
        // r = A x - b (left: preconditioned); V_0 = -r / |r|.  On the first cycle of a
        // right-preconditioned solve the caller may have formed A x - b in t already (r0_in_t).
        if (!(right && r0_in_t)) {
            RC(k_saddle_multiply(ctx, pos, x, t));
            RC(k_axpby(ctx, n, 1.0f, t, -1.0f, b, t));
        }
        r0_in_t = false;
        float *r = t;
        if (!right) {
            RC(k_saddle_precond(ctx, t, w));
            r = w;
        }
        RC(k_nrm2(ctx, n, r, d_h));
        RC(hostsync(1));
        const double beta = std::sqrt(ctx->h_part[0]);
        resid = beta;
        if (st && iters == 0 && restarts == 0) st->initial_residual = (float)(beta / bn);
        if (resid <= target || iters >= maxit || beta == 0.0) break;
        RC(k_axpby(ctx, n, (float)(-1.0 / beta), r, 0.0f, nullptr, V));
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        // Arnoldi step j runs entirely on the device (its new basis vector is normalised there), and
        // its Hessenberg column is read back into slot j & 1 of the pinned buffer; so step j + 1 can
        // be enqueued before the host has that column, and the host waits only for it.  Near
        // convergence (estimate within 1.5 x the target) the next step is enqueued after the check.
        auto enqueue = [&](int j) -> int {
            // the new direction w is formed in place of the next basis vector
            float *vj = V + (int64_t)j * n, *vn = V + (int64_t)(j + 1) * n;
            if (right) {
                float *zj = Z + (int64_t)j * n;
                RC(k_saddle_precond(ctx, vj, zj));
                RC(k_saddle_multiply(ctx, pos, zj, vn));
            } else {
                RC(k_saddle_multiply(ctx, pos, vj, t));
                RC(k_saddle_precond(ctx, t, vn));
            }
            hipEvent_t ev;
            prof_begin(ctx, PROF_GMRES_VEC, &ev);
            // CGS2: two passes of h = V^T w, w -= V h.  The second pass also forms |w'|^2 (w' is
            // the next basis slot itself); after it |w''|^2 = |w'|^2 - |h2|^2, with h2 tiny, so
            // the norm needs no pass of its own
            RC(k_multidot(ctx, n, j + 1, V, n, vn, d_h));
            RC(k_multi_axpy(ctx, n, j + 1, V, n, d_h, vn));
            if (ctx->p.gs_passes == 1) {   // one classical pass: no second projection, direct norm
                RC(k_nrm2(ctx, n, vn, d_h + KM + j + 1));
                RC(k_scale_by_norm(ctx, n, vn, d_h + KM + j + 1, 0, d_h + 2 * KM + 1));
            } else {
                RC(k_multidot(ctx, n, j + 2, V, n, vn, d_h + KM));
                RC(k_multi_axpy(ctx, n, j + 1, V, n, d_h + KM, vn));
                RC(k_scale_by_norm(ctx, n, vn, d_h + KM, j + 1, d_h + 2 * KM + 1));
            }
            prof_end(ctx, PROF_GMRES_VEC, ev);
            CK(hipMemcpyAsync(ctx->h_part + (size_t)(j & 1) * HS, d_h, sizeof(double) * HS, hipMemcpyDeviceToHost,
                              ctx->stream));
            CK(hipEventRecord(ctx->gm_ev[j & 1], ctx->stream));
            return 0;
        };
        RC(enqueue(0));
        int i = -1, queued = 0;
        do {
            ++i;
            ++iters;
            if (queued == i && i + 1 < R && iters < maxit && resid > 1.5 * target) {
                RC(enqueue(i + 1));
                queued = i + 1;
            }
            CK(hipEventSynchronize(ctx->gm_ev[i & 1]));
            const double *hs = ctx->h_part + (size_t)(i & 1) * HS;
            const bool two_pass = ctx->p.gs_passes != 1;
            for (int k = 0; k <= i; ++k) H[(size_t)k * R + i] = two_pass ? hs[k] + hs[KM + k] : hs[k];
            H[(size_t)(i + 1) * R + i] = hs[2 * KM + 1];
            for (int k = 0; k < i; ++k) {
                const double a = H[(size_t)k * R + i], c = H[(size_t)(k + 1) * R + i];
                H[(size_t)k * R + i] = cs[k] * a + sn[k] * c;
                H[(size_t)(k + 1) * R + i] = -sn[k] * a + cs[k] * c;
            }
            const double a = H[(size_t)i * R + i], c = H[(size_t)(i + 1) * R + i];
            if (c == 0.0) { cs[i] = 1.0; sn[i] = 0.0; }
            else if (std::fabs(c) > std::fabs(a)) { double tt = a / c; sn[i] = 1.0 / std::sqrt(1.0 + tt * tt); cs[i] = tt * sn[i]; }
            else { double tt = c / a; cs[i] = 1.0 / std::sqrt(1.0 + tt * tt); sn[i] = tt * cs[i]; }
            H[(size_t)i * R + i] = cs[i] * a + sn[i] * c;
            H[(size_t)(i + 1) * R + i] = 0.0;
            g[i + 1] = -sn[i] * g[i];
            g[i] = cs[i] * g[i];
            resid = std::fabs(g[i + 1]);
            if (!(resid <= target) && iters < maxit && i < R - 1 && queued == i) {
                RC(enqueue(i + 1));
                queued = i + 1;
            }
        } while (!(resid <= target) && iters < maxit && i < R - 1);
        // y = H^-1 g ; x += V y (left) or Z y (right)
        for (int j = i; j >= 0; --j) {
            double s = g[j];
            for (int k = j + 1; k <= i; ++k) s -= H[(size_t)j * R + k] * yv[k];
            yv[j] = s / H[(size_t)j * R + j];
        }
        double *coef = ctx->h_part + 2 * (size_t)HS;
        for (int j = 0; j <= i; ++j) coef[j] = -yv[j];
        CK(hipMemcpyAsync(d_h, coef, sizeof(double) * (i + 1), hipMemcpyHostToDevice, ctx->stream));
        RC(k_multi_axpy(ctx, n, i + 1, right ? Z : V, n, d_h, x));
        CK(hipStreamSynchronize(ctx->stream));
        ++restarts;
        if (resid <= target || iters >= maxit) break;
    }
    if (st) {
        st->iterations = iters;
        st->restarts = restarts;
        st->residual = (float)(resid / (bn > 0 ? bn : 1.0));
    }
    return (resid <= target || bn == 0.0) ? 0 : FSD_W_NOT_CONVERGED;
}

// The saddle solve with this step's preconditioner.  With the learned factor
// the solve may take at most IC(0)'s iteration count for the same solve
// (which: 0 the step's solve, 1 and 2 the two random-finite-difference solves
// of a Brownian step) at the last calibration step, which records it; a solve
// that reaches it without converging continues from its iterate with IC(0),
// built then (the iteration fallback), and IC(0) stays for the rest of the step.
static int solve_saddle(fsd_ctx *ctx, const float4 *pos, const float *b, float *x, int maxit, fsd_solve_stats *st,
                        bool r0_in_t, int which = 0)
{
    const bool learned = ctx->precond_mode == 2 && lp_active(ctx);
    int cap = maxit;
    if (learned && ctx->lp_iter_cap[which] > 0) cap = std::min(maxit, ctx->lp_iter_cap[which]);
    int rc = gmres(ctx, pos, b, x, ctx->p.gmres_tol, cap, ctx->p.gmres_restart, st, r0_in_t);
    if (rc < 0) return rc;
    if (ctx->lp_force_ichol) ctx->lp_iter_cap[which] = st->iterations;
    st->precond_used = learned ? 1 : 0;
    st->fallback = ctx->lp_step_fallback;
    if (learned && rc == FSD_W_NOT_CONVERGED && cap < maxit) {
        ctx->lp_iter_fallbacks++;
        ctx->lp_step_fallback = 2;
        RC(precond_ichol(ctx));
        fsd_solve_stats s2;
        memset(&s2, 0, sizeof s2);
        s2.true_residual = -1.f;
        rc = gmres(ctx, pos, b, x, ctx->p.gmres_tol, maxit - st->iterations, ctx->p.gmres_restart, &s2, false);
        if (rc < 0) return rc;
        st->iterations += s2.iterations;
        st->restarts += s2.restarts;
        st->residual = s2.residual;
        st->true_residual = s2.true_residual;
        st->precond_used = 0;
        st->fallback = 2;
    }
    return rc;
}

// ----------------------------------------------------------- Brownian motion
// Noise streams of one step: far-field real space (11N), far-field grid (6M),
// near field (6N), random finite difference (6N).
enum { BRO_FF_REAL = 0, BRO_FF_GRID = 1, BRO_NF = 2, BRO_RFD = 3 };

static int alloc_brownian(fsd_ctx *ctx)
{
    if (ctx->bro_alloc) return 0;
    const int64_t N = ctx->N, n11 = 11 * N;
    CK(hipMalloc(&ctx->d_lanV, sizeof(float) * (BRO_MMAX + 1) * n11));
    CK(hipMalloc(&ctx->d_lanw, sizeof(float) * n11));
    CK(hipMalloc(&ctx->d_lany0, sizeof(float) * n11));
    CK(hipMalloc(&ctx->d_lan_h, sizeof(double) * 256));
    CK(hipHostMalloc((void **)&ctx->h_lan, sizeof(double) * 256, hipHostMallocDefault));
    CK(hipMalloc(&ctx->d_noise, sizeof(float) * (n11 + 6 * (int64_t)ctx->M + 12 * N)));
    CK(hipMalloc(&ctx->d_posp, sizeof(float4) * N));
    CK(hipMalloc(&ctx->d_rfd_rhs, sizeof(float) * 17 * N));
    CK(hipMalloc(&ctx->d_rfd_sol, sizeof(float) * 17 * N));
    ctx->bro_alloc = true;
    return 0;
}

// unit normal noise of one stream: the caller's injected buffer, or Philox
// normals keyed by (seed, timestep, stream)
static int bro_noise(fsd_ctx *ctx, int stream, uint32_t timestep, const float **out)
{
    const int64_t N = ctx->N, n11 = 11 * N, nG = 6 * (int64_t)ctx->M;
    const float *inj[4] = {ctx->inj_ff_real, ctx->inj_ff_grid, ctx->inj_nf, ctx->inj_rfd};
    const int64_t off[4] = {0, n11, n11 + nG, n11 + nG + 6 * N};
    const int64_t len[4] = {n11, nG, 6 * N, 6 * N};
    if (inj[stream]) { *out = inj[stream]; return 0; }
    float *dst = ctx->d_noise + off[stream];
    RC(k_normals(ctx, len[stream], timestep, (uint32_t)stream, 1.0f, dst));
    *out = dst;
    return 0;
}

// y = |psi| V_m T_m^(1/2) e_1, Brownian_FarField_Chow_Saad /
// Brownian_NearField_Chow_Saad (Brownian_FarField.cu:597-808,
// Brownian_NearField.cu:174-389) after Chow and Saad (SIAM J. Sci. Comput.
// 2014): the Krylov size starts at *m_io (the previous step's) and then grows
// one vector at a time until the relative change of y drops below tol or the
// basis holds BRO_MMAX - 1 vectors.  Each new vector is reorthogonalised
// against the whole basis (two classical Gram-Schmidt passes with FP64 dot
// products); T_m^(1/2) e_1 is formed in FP64 on the host.
// op: 0 = real-space far-field mobility (11N), 1 = lubrication R_FU (6N).
static int lanczos_sqrt(fsd_ctx *ctx, const float4 *pos, int op, const float *psi, float *y, int *m_io, float tol)
{
    const int64_t n = (op == 0 ? 11 : 6) * (int64_t)ctx->N;
    float *V = ctx->d_lanV, *w = ctx->d_lanw, *y0 = ctx->d_lany0;
    double *dh = ctx->d_lan_h, *hh = ctx->h_lan;
    const double tol_beta = 1e-8;
    auto fetch = [&](int cnt) -> int {
        CK(hipMemcpyAsync(hh, dh, sizeof(double) * cnt, hipMemcpyDeviceToHost, ctx->stream));
        CK(hipStreamSynchronize(ctx->stream));
        return 0;
    };
    auto apply = [&](const float *x, float *out) -> int {
        if (op == 0) return k_mobility(ctx, pos, x, out, 1, nullptr);
        return k_rfu_apply(ctx, x, out, nullptr, 1.0f);
    };
    RC(k_nrm2(ctx, n, psi, dh));
    RC(fetch(1));
    const double xnorm = std::sqrt(hh[0]);
    if (xnorm == 0.0) {
        RC(k_fill(ctx, n, 0.f, y));
        return 0;
    }
    RC(k_axpby(ctx, n, (float)(1.0 / xnorm), psi, 0.0f, nullptr, V));
    std::vector<double> alpha, off;
    bool breakdown = false;
    // appends alpha_j and (unless the Krylov space is exhausted) off_j and v_{j+1}
    auto step = [&](int j) -> int {
        float *vj = V + (int64_t)j * n;
        RC(apply(vj, w));
        double a = 0.0;
        for (int pass = 0; pass < 2; ++pass)
            for (int c0 = 0; c0 <= j; c0 += 64) {
                const int K = std::min(64, j + 1 - c0);
                RC(k_multidot(ctx, n, K, V + (int64_t)c0 * n, n, w, dh));
                RC(k_multi_axpy(ctx, n, K, V + (int64_t)c0 * n, n, dh, w));
                if (j >= c0 && j < c0 + K) {
                    RC(fetch(K));
                    a += hh[j - c0];
                }
            }
        RC(k_nrm2(ctx, n, w, dh));
        RC(fetch(1));
        const double b = std::sqrt(hh[0]);
        alpha.push_back(a);
        if (b < tol_beta) { breakdown = true; return 0; }
        off.push_back(b);
        RC(k_axpby(ctx, n, (float)(1.0 / b), w, 0.0f, nullptr, V + (int64_t)(j + 1) * n));
        return 0;
    };
    // out = xnorm V_m T_m^(1/2) e_1
    auto combine = [&](int m, float *out) -> int {
        std::vector<double> t(m);
        tridiag_sqrt_e1(m, alpha.data(), off.data(), t.data());
        RC(k_fill(ctx, n, 0.f, out));
        for (int c0 = 0; c0 < m; c0 += 64) {
            const int K = std::min(64, m - c0);
            for (int k = 0; k < K; ++k) hh[k] = -xnorm * t[c0 + k];
            CK(hipMemcpyAsync(dh, hh, sizeof(double) * K, hipMemcpyHostToDevice, ctx->stream));
            RC(k_multi_axpy(ctx, n, K, V + (int64_t)c0 * n, n, dh, out));
            CK(hipStreamSynchronize(ctx->stream));
        }
        return 0;
    };
    int m = std::max(1, std::min(*m_io, BRO_MMAX - 2));
    for (int j = 0; j < m; ++j) {
        RC(step(j));
        if (breakdown) { m = j + 1; break; }
    }
    RC(combine(m, y0));
    double err = 1.0;
    bool wrote = false;
    while (!breakdown && err > tol && m < BRO_MMAX - 1) {
        RC(step(m));
        RC(combine(m + 1, y));
        wrote = true;
        RC(k_axpby(ctx, n, 1.0f, y0, -1.0f, y, y0));      // y0 <- y0 - y
        RC(k_nrm2(ctx, n, y0, dh));
        RC(k_nrm2(ctx, n, y, dh + 1));
        RC(fetch(2));
        err = std::sqrt(hh[0]) / std::max(std::sqrt(hh[1]), 1e-300);
        RC(k_copy(ctx, n, y, y0));
        ++m;
    }
    if (!wrote) RC(k_copy(ctx, n, y0, y));
    *m_io = m - 1;
    return 0;
}

// far-field Brownian slip U_B^ff (11N) into out: wave space from the grid
// noise plus sqrt(2T/dt) M_real^(1/2) psi (Brownian_FarField_SlipVelocity,
// Brownian_FarField.cu:829-1061)
static int bro_farfield(fsd_ctx *ctx, const float4 *pos, uint32_t timestep, float *out)
{
    const float T = ctx->p.T, dt = ctx->p.dt;
    const fsd_pse_info &ps = ctx->pse;
    const float *psi = nullptr, *grid_noise = nullptr;
    RC(bro_noise(ctx, BRO_FF_REAL, timestep, &psi));
    RC(bro_noise(ctx, BRO_FF_GRID, timestep, &grid_noise));
    RC(lanczos_sqrt(ctx, pos, 0, psi, out, &ctx->m_lanczos_ff, ctx->p.error));
    ctx->last_lanczos_ff = ctx->m_lanczos_ff;
    RC(k_axpby(ctx, 11 * (int64_t)ctx->N, std::sqrt(2.0f * T / dt), out, 0.0f, nullptr, out));
    const float quadW = ps.gridh_x * ps.gridh_y * ps.gridh_z;
    RC(k_ff_noise(ctx, grid_noise, std::sqrt(T / dt / quadW)));
    RC(k_wave_contract(ctx, pos));
    RC(k_wave_to_general(ctx, out));
    return 0;
}

// near-field Brownian force F_B^nf = R_FU^(1/2) psi, psi ~ N(0, 2T/dt) (6N)
// (Brownian_NearField_Force, Brownian_NearField.cu:407-464)
static int bro_nearfield(fsd_ctx *ctx, const float4 *pos, uint32_t timestep, float *out)
{
    const float *psi = nullptr;
    RC(bro_noise(ctx, BRO_NF, timestep, &psi));
    RC(lanczos_sqrt(ctx, pos, 1, psi, out, &ctx->m_lanczos_nf, ctx->p.error));
    ctx->last_lanczos_nf = ctx->m_lanczos_nf;
    RC(k_axpby(ctx, 6 * (int64_t)ctx->N, std::sqrt(2.0f * ctx->p.T / ctx->p.dt), out, 0.0f, nullptr, out));
    return 0;
}

// Integrator_RFD (Integrator.cu:549-718): drift (T / eps)(U+ - U-) (6N) with
// U+- the velocities of saddle solves for the right-hand side [0; 0; -psi] at
// positions wrap(x +- eps psi / 2), keeping this step's neighbour list and
// preconditioner; the second solve starts from the first's solution.  The
// lubrication pairs and the spreading set-up are rebuilt for each shifted
// configuration and restored for x afterwards.
static int bro_rfd(fsd_ctx *ctx, const float4 *pos, uint32_t timestep, float *drift)
{
    const int64_t N = ctx->N, n17 = 17 * N;
    const float eps = ctx->p.error;
    const float *psi = nullptr;
    RC(bro_noise(ctx, BRO_RFD, timestep, &psi));
    RC(k_fill(ctx, n17, 0.f, ctx->d_rfd_rhs));
    RC(k_fill(ctx, n17, 0.f, ctx->d_rfd_sol));
    RC(k_axpby(ctx, 6 * N, -1.0f, psi, 0.0f, nullptr, ctx->d_rfd_rhs + 11 * N));
    const int maxit = ctx->p.gmres_maxiter;
    for (int side = 0; side < 2; ++side) {
        RC(k_shift_wrap(ctx, pos, psi, side == 0 ? 0.5f * eps : -0.5f * eps, ctx->d_posp));
        RC(k_build_lub_pairs(ctx, ctx->d_posp));
        ctx->spread_valid = false;
        fsd_solve_stats st;
        memset(&st, 0, sizeof st);
        st.true_residual = -1.f;
        const int rc = solve_saddle(ctx, ctx->d_posp, ctx->d_rfd_rhs, ctx->d_rfd_sol, maxit, &st, false, 1 + side);
        if (rc < 0) return rc;
        if (side == 0) RC(k_copy(ctx, 6 * N, ctx->d_rfd_sol + 11 * N, drift));
    }
    const float fac = ctx->p.T / eps;
    RC(k_axpby(ctx, 6 * N, fac, drift, -fac, ctx->d_rfd_sol + 11 * N, drift));
    RC(k_build_lub_pairs(ctx, pos));
    ctx->spread_valid = false;
    RC(k_setup_spread(ctx, pos));
    return 0;
}

static int true_residual(fsd_ctx *ctx, const float4 *pos, const float *b, const float *x, float *out)
{
    const int64_t n = 17 * (int64_t)ctx->N;
    double *d_h = ctx->d_part + (int64_t)ctx->part_blocks * ctx->kmax;
    RC(k_saddle_multiply(ctx, pos, x, ctx->d_t));
    RC(k_axpby(ctx, n, 1.0f, b, -1.0f, ctx->d_t, ctx->d_t));
    RC(k_nrm2(ctx, n, ctx->d_t, d_h));
    RC(k_nrm2(ctx, n, b, d_h + 1));
    CK(hipMemcpyAsync(ctx->h_part, d_h, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    *out = (float)(std::sqrt(ctx->h_part[0]) / std::max(std::sqrt(ctx->h_part[1]), 1e-300));
    return 0;
}

// ------------------------------------------------------------ C ABI
extern "C" {

int fsd_create(const fsd_params *p, fsd_ctx **out)
{
    *out = nullptr;
    fsd_ctx *ctx = new fsd_ctx();
    ctx->p = *p;
    fsd_params &q = ctx->p;
    if (q.gmres_tol <= 0.f) q.gmres_tol = q.error;
    if (q.gmres_restart <= 0) q.gmres_restart = 50;
    if (q.gmres_restart > 255) q.gmres_restart = 255;
    if (q.gmres_maxiter <= 0) q.gmres_maxiter = 1000;
    if (q.rlub <= 0.f) q.rlub = 4.0f;
    if (q.rp <= 0.f) q.rp = 2.1f;
    if (q.r_buff < 0.f) q.r_buff = 0.4f;
    if (q.gs_passes <= 0) q.gs_passes = 2;
    if (q.gs_passes > 2) q.gs_passes = 2;
    if (q.jacobi_iters <= 0) q.jacobi_iters = 3;
    if (q.factor_sweeps < 0) q.factor_sweeps = 0;
    if (q.precond < 0 || q.precond > 2) {
        ctx->err = "precond must be 0 (IC(0), exact solves), 1 (IC(0), Jacobi) or 2 (learned, Jacobi)";
        *out = ctx;
        return FSD_E_PARAM;
    }
    if (q.fallback_jacobi < 0) q.fallback_jacobi = 0;
    ctx->precond_mode = q.precond;
    ctx->jac_k = q.jacobi_iters;
    ctx->ic_jac_k = q.precond == 1 ? q.jacobi_iters : (q.precond == 2 ? q.fallback_jacobi : 0);
    if (q.N <= 0 || q.xi <= 0.f || q.error <= 0.f || q.error >= 1.f || q.Lx <= 0.f || q.Ly <= 0.f || q.Lz <= 0.f) {
        ctx->err = "invalid fsd_params";
        *out = ctx;
        return FSD_E_PARAM;
    }
    ctx->N = q.N;
    set_box(ctx, q.Lx, q.Ly, q.Lz, q.xy, q.xz, q.yz);
    int rc = pse_params(ctx);
    if (rc) { *out = ctx; return rc; }
    cell_dims(ctx);
    // the HIP runtime is initialised on the current device before anything else (this may be
    // the process's first HIP call), then rocFFT once per process
    {
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipSetDevice(dev);
        if (e == hipSuccess) e = hipFree(nullptr);
        if (e != hipSuccess) {
            ctx->err = std::string("HIP runtime initialisation failed: ") + hipGetErrorString(e);
            *out = ctx;
            return FSD_E_HIP;
        }
    }
    static bool fft_ready = false;
    if (!fft_ready) { rocfft_setup(); fft_ready = true; }
    {
        const hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            ctx->err = std::string("hipStreamCreate failed: ") + hipGetErrorString(e);
            *out = ctx;
            return FSD_E_HIP;
        }
    }
    // device address of the allocation-free entry point's scratch array (once per device)
    if (noalloc_init_device() != 0) {
        ctx->err = "hipGetSymbolAddress of the scratch array failed";
        *out = ctx;
        return FSD_E_HIP;
    }
    ctx->slots.resize(PROF_COUNT);
    for (int s = 0; s < PROF_COUNT; ++s) { ctx->slots[s].name = kProfNames[s]; ctx->slots[s].launches = 0; ctx->slots[s].ms = 0; }
    rc = alloc_all(ctx);
    *out = ctx;
    if (rc) return rc;
    return k_gridk(ctx) ? FSD_E_HIP : 0;
}

void fsd_destroy(fsd_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    free_all(ctx);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *fsd_last_error(const fsd_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int fsd_get_pse_info(const fsd_ctx *ctx, fsd_pse_info *out)
{
    *out = ctx->pse;
    return 0;
}

int fsd_get_ewald_table(const fsd_ctx *ctx, float *h_tab, int32_t capacity)
{
    if ((size_t)capacity < ctx->h_ewald.size()) return FSD_E_CAPACITY;
    memcpy(h_tab, ctx->h_ewald.data(), sizeof(float) * ctx->h_ewald.size());
    return 0;
}

int fsd_set_resistance_table(fsd_ctx *ctx, const float *h_dist, const float *h_vals)
{
    ctx->h_res_dist.assign(h_dist, h_dist + 1000);
    ctx->h_res_vals.assign(h_vals, h_vals + 22000);
    CK(hipMemcpy(ctx->d_res_dist, h_dist, sizeof(float) * 1000, hipMemcpyHostToDevice));
    CK(hipMemcpy(ctx->d_res_vals, h_vals, sizeof(float) * 22000, hipMemcpyHostToDevice));
    ctx->have_table = true;
    ctx->lub_valid = false;
    ctx->precond_valid = false;
    return 0;
}

// Stokes::setFriction (Stokes.cc:1132-1191).  TYPE1 adds, for gaps h <= h0,
// alpha_k (2 h^2 / h0^3 - 3 h / h0^2 + 1 / h) to six tangential resistance
// functions (zero at h = h0).  The reference's comments name YA11, YA12,
// YB11, YB12, YC11, YC12 (columns 2, 3, 4, 5, 8, 9) but its code writes
// columns 3, 4, 5, 6, 9, 10; parity = 1 keeps the columns the code writes.
int fsd_set_friction(fsd_ctx *ctx, const char *type, float h0, const float *h_alpha6)
{
    std::string t = type ? type : "NONE";
    for (auto &c : t) c = (char)toupper(c);
    if (t == "NONE") return 0;
    if (t != "TYPE1") { ctx->err = "Invalid Friction Type. Allowable types are: None, Type1"; return FSD_E_PARAM; }
    if (!ctx->have_table) { ctx->err = "setFriction must follow setResistanceTable"; return FSD_E_STATE; }
    static const int cols_code[6] = {3, 4, 5, 6, 9, 10}, cols_named[6] = {2, 3, 4, 5, 8, 9};
    static const float sgn[6] = {1.f, -1.f, -1.f, 1.f, 1.f, 1.f};
    const int *cols = ctx->p.parity ? cols_code : cols_named;
    for (int ii = 0; ii < 1000; ++ii) {
        const float h = (float)((double)ctx->h_res_dist[ii] - 2.0);
        if (!(h <= h0)) continue;
        const float h2 = h * h, h02 = h0 * h0, h03 = h0 * h02;
        const double poly = 2.0 / h03 * h2 - 3.0 / h02 * h + 1.0 / h;
        for (int k = 0; k < 6; ++k) {
            const float add = (float)(h_alpha6[k] * poly);
            ctx->h_res_vals[22 * ii + cols[k]] += sgn[k] * add;
        }
    }
    CK(hipMemcpy(ctx->d_res_vals, ctx->h_res_vals.data(), sizeof(float) * 22000, hipMemcpyHostToDevice));
    ctx->lub_valid = false;
    ctx->precond_valid = false;
    return 0;
}

// the resistance table in use (after friction), h_vals[22000]
int fsd_get_resistance_table(const fsd_ctx *ctx, float *h_vals)
{
    if (!ctx->have_table) return FSD_E_STATE;
    std::copy(ctx->h_res_vals.begin(), ctx->h_res_vals.end(), h_vals);
    return 0;
}

int fsd_set_temperature(fsd_ctx *ctx, float T)
{
    ctx->p.T = T;
    return 0;
}

int fsd_set_box(fsd_ctx *ctx, float Lx, float Ly, float Lz, float xy, float xz, float yz)
{
    if (Lx != ctx->box.Lx || Ly != ctx->box.Ly || Lz != ctx->box.Lz) {
        ctx->err = "box lengths are fixed at creation; only the tilt may change";
        return FSD_E_PARAM;
    }
    set_box(ctx, Lx, Ly, Lz, xy, xz, yz);
    ctx->nlist_valid = false;
    ctx->spread_valid = false;
    return k_gridk(ctx) ? FSD_E_HIP : 0;
}

int fsd_build_nlist(fsd_ctx *ctx, const float *d_pos4)
{
    int rc = k_build_nlist(ctx, reinterpret_cast<const float4 *>(d_pos4));
    if (rc) return rc;
    ctx->nlist_builds++;
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int fsd_get_nlist(fsd_ctx *ctx, int32_t *h_nneigh, int32_t *h_nbrs, int64_t capacity, int64_t *total)
{
    if (!ctx->nlist_valid) { ctx->err = "neighbour list not built"; return FSD_E_STATE; }
    const int N = ctx->N;
    std::vector<int> nn(N), lst((size_t)N * ctx->maxn);
    CK(hipMemcpy(nn.data(), ctx->d_nneigh, sizeof(int) * N, hipMemcpyDeviceToHost));
    CK(hipMemcpy(lst.data(), ctx->d_nlist, sizeof(int) * lst.size(), hipMemcpyDeviceToHost));
    int64_t tot = 0;
    for (int i = 0; i < N; ++i) {
        std::vector<int> v(lst.begin() + (size_t)i * ctx->maxn, lst.begin() + (size_t)i * ctx->maxn + nn[i]);
        std::sort(v.begin(), v.end());
        if (h_nneigh) h_nneigh[i] = nn[i];
        for (int j : v) {
            if (h_nbrs && tot < capacity) h_nbrs[tot] = j;
            ++tot;
        }
    }
    *total = tot;
    return tot > capacity && h_nbrs ? FSD_E_CAPACITY : 0;
}

// The lubrication pair records are a function of the positions handed in, and an
// entry point cannot tell whether the caller moved the particles since the last
// build (fsd_step moves them; callers update positions in place), so every entry
// point that applies R_FU builds them for its own positions.
// An operator entry point may be handed positions that moved since the list was
// built (fsd_step moves them): the list is rebuilt, as at the start of a step, when
// some particle has moved r_buff / 2 since its build (HOOMD's distance check).
static int current_nlist(fsd_ctx *ctx, const float4 *pos)
{
    if (!ctx->nlist_valid) { ctx->err = "neighbour list not built"; return FSD_E_STATE; }
    bool rebuilt = false;
    RC(k_update_nlist(ctx, pos, &rebuilt));
    ctx->nlist_builds += rebuilt;
    return 0;
}

static int ensure_lub(fsd_ctx *ctx, const float4 *pos)
{
    RC(current_nlist(ctx, pos));
    RC(k_build_lub_pairs(ctx, pos));
    return 0;
}

int fsd_set_noise(fsd_ctx *ctx, const float *d_ff_real11, const float *d_ff_grid6M, const float *d_nf6,
                  const float *d_rfd6)
{
    ctx->inj_ff_real = d_ff_real11;
    ctx->inj_ff_grid = d_ff_grid6M;
    ctx->inj_nf = d_nf6;
    ctx->inj_rfd = d_rfd6;
    return 0;
}

int fsd_brownian_farfield(fsd_ctx *ctx, const float *d_pos4, uint32_t timestep, float *d_out11)
{
    const float4 *pos = reinterpret_cast<const float4 *>(d_pos4);
    RC(current_nlist(ctx, pos));
    RC(alloc_brownian(ctx));
    RC(k_gridk(ctx));
    ctx->spread_valid = false;
    RC(k_setup_spread(ctx, pos));
    RC(bro_farfield(ctx, pos, timestep, d_out11));
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int fsd_brownian_nearfield(fsd_ctx *ctx, const float *d_pos4, uint32_t timestep, float *d_out6)
{
    const float4 *pos = reinterpret_cast<const float4 *>(d_pos4);
    RC(ensure_lub(ctx, pos));
    RC(alloc_brownian(ctx));
    RC(bro_nearfield(ctx, pos, timestep, d_out6));
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int fsd_rfd_drift(fsd_ctx *ctx, const float *d_pos4, uint32_t timestep, float *d_out6)
{
    const float4 *pos = reinterpret_cast<const float4 *>(d_pos4);
    if (!ctx->precond_valid) { ctx->err = "preconditioner not built"; return FSD_E_STATE; }
    RC(alloc_brownian(ctx));
    RC(k_gridk(ctx));
    RC(bro_rfd(ctx, pos, timestep, d_out6));
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int fsd_rfu(fsd_ctx *ctx, const float *d_pos4, const float *d_U6, float *d_F6)
{
    const float4 *pos = reinterpret_cast<const float4 *>(d_pos4);
    RC(ensure_lub(ctx, pos));
    RC(k_rfu_apply(ctx, d_U6, d_F6, nullptr, 1.0f));
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int fsd_rfe(fsd_ctx *ctx, const float *d_pos4, float shear_rate, const float *d_ori3, float *d_F6)
{
    if (!ctx->have_table) { ctx->err = "resistance table not set"; return FSD_E_STATE; }
    RC(current_nlist(ctx, reinterpret_cast<const float4 *>(d_pos4)));
    RC(k_rfe(ctx, reinterpret_cast<const float4 *>(d_pos4), shear_rate, d_ori3, d_F6));
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int fsd_rsu(fsd_ctx *ctx, const float *d_pos4, const float *d_U6, float *d_S5)
{
    if (!ctx->have_table) { ctx->err = "resistance table not set"; return FSD_E_STATE; }
    RC(current_nlist(ctx, reinterpret_cast<const float4 *>(d_pos4)));
    RC(k_rsu(ctx, reinterpret_cast<const float4 *>(d_pos4), d_U6, d_S5));
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int fsd_rse(fsd_ctx *ctx, const float *d_pos4, float shear_rate, const float *d_ori3, float *d_S5)
{
    if (!ctx->have_table) { ctx->err = "resistance table not set"; return FSD_E_STATE; }
    RC(current_nlist(ctx, reinterpret_cast<const float4 *>(d_pos4)));
    RC(k_rse(ctx, reinterpret_cast<const float4 *>(d_pos4), shear_rate, d_ori3, d_S5));
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int fsd_mobility(fsd_ctx *ctx, const float *d_pos4, const float *d_gF11, float *d_gU11, int32_t parts)
{
    RC(current_nlist(ctx, reinterpret_cast<const float4 *>(d_pos4)));
    ctx->spread_valid = false;
    RC(k_mobility(ctx, reinterpret_cast<const float4 *>(d_pos4), d_gF11, d_gU11, parts, nullptr));
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int fsd_autotune(fsd_ctx *ctx, const float *d_pos4, int32_t *h_choice8, float *h_us3)
{
    const float4 *pos = reinterpret_cast<const float4 *>(d_pos4);
    RC(ensure_lub(ctx, pos));
    ctx->spread_valid = false;
    ctx->tune = fsd_tuning();
    RC(autotune_run(ctx, pos));
    return fsd_get_tuning(ctx, h_choice8, h_us3);
}

int fsd_get_tuning(fsd_ctx *ctx, int32_t *h_choice8, float *h_us3)
{
    const fsd_tuning &t = ctx->tune;
    if (h_choice8) {
        const int32_t c[8] = {t.spread, t.contract, t.xpass, t.real_g, t.real_blk, t.rfu_g, t.rfu_blk, t.tuned};
        for (int i = 0; i < 8; ++i) h_choice8[i] = c[i];
    }
    if (h_us3)
        for (int i = 0; i < 3; ++i) h_us3[i] = t.us[i];
    return 0;
}

int fsd_set_tuning(fsd_ctx *ctx, const int32_t *h_choice7)
{
    const int32_t *c = h_choice7;
    const bool ok = (c[0] == 0 || c[0] == 1) && (c[1] == 0 || c[1] == 1) && (c[2] == 0 || c[2] == 1) &&
                    (c[3] == 4 || c[3] == 8 || c[3] == 16) && (c[4] == 128 || c[4] == 256) &&
                    (c[5] == 8 || c[5] == 16) && (c[6] == 128 || c[6] == 256);
    if (!ok) {
        ctx->err = "tuning: spread, contract, xpass in {0, 1}; real lanes in {4, 8, 16}; R_FU lanes in {8, 16}; "
                   "workgroups of 128 or 256";
        return FSD_E_PARAM;
    }
    fsd_tuning &t = ctx->tune;
    t.spread = c[0]; t.contract = c[1]; t.xpass = c[2];
    t.real_g = c[3]; t.real_blk = c[4]; t.rfu_g = c[5]; t.rfu_blk = c[6];
    t.tuned = 1;
    return 0;
}

int fsd_build_precond(fsd_ctx *ctx, const float *d_pos4, fsd_solve_stats *stats)
{
    RC(current_nlist(ctx, reinterpret_cast<const float4 *>(d_pos4)));
    RC(build_precond(ctx, reinterpret_cast<const float4 *>(d_pos4)));
    if (stats) {
        stats->ichol_relaxer = ctx->ichol_relaxer;
        stats->ichol_attempts = ctx->ichol_attempts;
        stats->nnz = ctx->pch.nnz_ref;
        stats->lower_blocks = ctx->pch.nlow;
        stats->levels = ctx->pch.nlev_f;
        stats->sweep_steps = ctx->pch.tp_ok ? ctx->pch.tp_f.nsteps + ctx->pch.tp_b.nsteps : 0;
        stats->precond_used = lp_active(ctx) ? 1 : 0;
        stats->fallback = ctx->lp_step_fallback;
    }
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int fsd_precond_apply(fsd_ctx *ctx, const float *d_x6, float *d_y6)
{
    if (!ctx->precond_valid) { ctx->err = "preconditioner not built"; return FSD_E_STATE; }
    if (ctx->precond_mode == 2 && lp_active(ctx)) RC(lp_apply(ctx, d_x6, nullptr, nullptr, d_y6, 0));
    else if (ctx->ic_jac_k > 0) RC(k_precond_apply_jacobi(ctx, d_x6, d_y6));
    else RC(k_precond_apply2(ctx, d_x6, d_x6, d_y6, d_y6, 1));
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// Reference-format export (pattern of Precondition_RFU_kernel + coo2csr and of
// cusolverSpXcsrpermHost; values of the factor on the permuted pattern).
int fsd_get_precond(fsd_ctx *ctx, int32_t *h_pn, int32_t *h_plist, int64_t plist_cap, int32_t *h_perm,
                    int32_t *h_rowptr, int32_t *h_col, float *h_val, int32_t *h_prow, int32_t *h_pcol, float *h_L,
                    int64_t nnz_cap, int64_t *nnz_out)
{
    if (!ctx->precond_valid) { ctx->err = "preconditioner not built"; return FSD_E_STATE; }
    const int N = ctx->N;
    const std::vector<int> &pn = ctx->h_pn, &perm = ctx->pch.perm;
    std::vector<int> phead(N + 1, 0);
    for (int i = 0; i < N; ++i) phead[i + 1] = phead[i] + pn[i];
    if (h_pn) memcpy(h_pn, pn.data(), sizeof(int) * N);
    if (h_plist) {
        if ((int64_t)phead[N] > plist_cap) return FSD_E_CAPACITY;
        for (int i = 0; i < N; ++i)
            memcpy(h_plist + phead[i], ctx->h_plstride.data() + (size_t)i * FSD_PSTRIDE, sizeof(int) * pn[i]);
    }
    if (h_perm) memcpy(h_perm, perm.data(), sizeof(int) * N);
    if (nnz_out) *nnz_out = ctx->pch.nnz_ref;
    if (!h_rowptr && !h_col && !h_val && !h_prow && !h_pcol && !h_L) return 0;
    if (ctx->pch.nnz_ref > nnz_cap) return FSD_E_CAPACITY;
    std::vector<int> iperm(N);
    for (int P = 0; P < N; ++P) iperm[perm[P]] = P;
    std::vector<int> lrow(N + 1), lcol(ctx->pch.nlow);
    std::vector<float> Ab(36 * (size_t)ctx->pch.nlow), Ad(36 * (size_t)N), Lb(36 * (size_t)ctx->pch.nlow), Ld(36 * (size_t)N);
    CK(hipMemcpy(lrow.data(), ctx->d_lrow, sizeof(int) * (N + 1), hipMemcpyDeviceToHost));
    if (ctx->pch.nlow) {
        CK(hipMemcpy(lcol.data(), ctx->d_lcol, sizeof(int) * ctx->pch.nlow, hipMemcpyDeviceToHost));
        CK(hipMemcpy(Ab.data(), ctx->d_Ablk, sizeof(float) * Ab.size(), hipMemcpyDeviceToHost));
        CK(hipMemcpy(Lb.data(), ctx->d_Lblk, sizeof(float) * Lb.size(), hipMemcpyDeviceToHost));
    }
    CK(hipMemcpy(Ad.data(), ctx->d_Adiag, sizeof(float) * Ad.size(), hipMemcpyDeviceToHost));
    CK(hipMemcpy(Ld.data(), ctx->d_Ldiag, sizeof(float) * Ld.size(), hipMemcpyDeviceToHost));
    // value of block (row particle i, column particle j), 6x6 element (a, c), original ids
    auto blockval = [&](const std::vector<float> &lowv, const std::vector<float> &diagv, int Pi, int Pj, int a,
                        int c) -> float {
        if (Pi == Pj) return diagv[(size_t)Pi * 36 + a * 6 + c];
        const int R0 = std::max(Pi, Pj), C0 = std::min(Pi, Pj);
        const int *b = lcol.data() + lrow[R0], *e = lcol.data() + lrow[R0 + 1];
        const int pos = (int)(std::lower_bound(b, e, C0) - lcol.data());
        const float *blk = lowv.data() + (size_t)pos * 36;
        return (Pi > Pj) ? blk[a * 6 + c] : blk[c * 6 + a];
    };
    // unpermuted CSR in the reference's COO order (rows ascending, neighbours ascending)
    if (h_rowptr || h_col || h_val) {
        int64_t k = 0;
        for (int i = 0; i < N; ++i) {
            const bool iso = pn[i] <= 1;
            for (int a = 0; a < 6; ++a) {
                const int row = 6 * i + a;
                if (h_rowptr) h_rowptr[row] = (int)k;
                if (iso) {
                    if (h_col) h_col[k] = row;
                    if (h_val) h_val[k] = 0.0f;
                    ++k;
                    continue;
                }
                for (int s = phead[i]; s < phead[i + 1]; ++s) {
                    const int j = ctx->h_plstride[(size_t)i * FSD_PSTRIDE + (s - phead[i])];
                    for (int c = 0; c < 6; ++c) {
                        if (h_col) h_col[k] = 6 * j + c;
                        if (h_val) h_val[k] = blockval(Ab, Ad, iperm[i], iperm[j], a, c);
                        ++k;
                    }
                }
            }
        }
        if (h_rowptr) h_rowptr[6 * N] = (int)k;
    }
    // permuted CSR (B = A(p,p), columns ascending) and the factor on that pattern
    if (h_prow || h_pcol || h_L) {
        int64_t k = 0;
        std::vector<int> cols;
        for (int P = 0; P < N; ++P) {
            const int i = perm[P];
            const bool iso = pn[i] <= 1;
            cols.clear();
            for (int s = 0; s < pn[i]; ++s) cols.push_back(iperm[ctx->h_plstride[(size_t)i * FSD_PSTRIDE + s]]);
            std::sort(cols.begin(), cols.end());
            for (int a = 0; a < 6; ++a) {
                const int row = 6 * P + a;
                if (h_prow) h_prow[row] = (int)k;
                if (iso) {
                    if (h_pcol) h_pcol[k] = row;
                    if (h_L) h_L[k] = Ld[(size_t)P * 36 + a * 6 + a];
                    ++k;
                    continue;
                }
                for (int Q : cols)
                    for (int c = 0; c < 6; ++c) {
                        const int col = 6 * Q + c;
                        if (h_pcol) h_pcol[k] = col;
                        if (h_L) h_L[k] = (col > row) ? 0.0f : blockval(Lb, Ld, P, Q, a, c);
                        ++k;
                    }
            }
        }
        if (h_prow) h_prow[6 * N] = (int)k;
    }
    return 0;
}

int fsd_saddle_multiply(fsd_ctx *ctx, const float *d_pos4, const float *d_x17, float *d_y17)
{
    const float4 *pos = reinterpret_cast<const float4 *>(d_pos4);
    RC(ensure_lub(ctx, pos));
    ctx->spread_valid = false;
    RC(k_saddle_multiply(ctx, pos, d_x17, d_y17));
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int fsd_saddle_precond(fsd_ctx *ctx, const float *d_b17, float *d_x17)
{
    if (!ctx->precond_valid) { ctx->err = "preconditioner not built"; return FSD_E_STATE; }
    const int64_t n = 17 * (int64_t)ctx->N;
    RC(k_copy(ctx, n, d_b17, ctx->d_t2));
    RC(k_saddle_precond(ctx, ctx->d_t2, d_x17));
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int fsd_saddle_solve(fsd_ctx *ctx, const float *d_pos4, const float *d_rhs17, float *d_sol17,
                     int32_t want_true_residual, fsd_solve_stats *stats)
{
    const float4 *pos = reinterpret_cast<const float4 *>(d_pos4);
    if (!ctx->precond_valid) { ctx->err = "preconditioner not built"; return FSD_E_STATE; }
    RC(ensure_lub(ctx, pos));
    ctx->spread_valid = false;
    fsd_solve_stats st;
    memset(&st, 0, sizeof st);
    st.true_residual = -1.f;
    int rc = solve_saddle(ctx, pos, d_rhs17, d_sol17, ctx->p.gmres_maxiter, &st, false);
    if (rc < 0) return rc;
    if (want_true_residual) RC(true_residual(ctx, pos, d_rhs17, d_sol17, &st.true_residual));
    st.ichol_relaxer = ctx->ichol_relaxer;
    st.ichol_attempts = ctx->ichol_attempts;
    st.nnz = ctx->pch.nnz_ref;
    st.nlist_total = ctx->nlist_total;
    st.lub_pairs = ctx->lub_total;
    st.lower_blocks = ctx->pch.nlow;
    st.levels = ctx->pch.nlev_f;
    st.sweep_steps = ctx->pch.tp_ok ? ctx->pch.tp_f.nsteps + ctx->pch.tp_b.nsteps : 0;
    if (stats) *stats = st;
    CK(hipStreamSynchronize(ctx->stream));
    return rc;
}

int fsd_force_law(fsd_ctx *ctx, const float *d_pos4, float T_ext, float *d_F6)
{
    RC(current_nlist(ctx, reinterpret_cast<const float4 *>(d_pos4)));
    RC(k_force_law(ctx, reinterpret_cast<const float4 *>(d_pos4), T_ext, d_F6));
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int fsd_euler_shear(fsd_ctx *ctx, float *d_pos4, float *d_ori3, float *d_pos_gb4, int32_t *d_image3,
                    const float *d_vel6, const float *d_mask1, const float *d_noise3, float dt, float shear_rate)
{
    RC(k_euler_shear(ctx, reinterpret_cast<float4 *>(d_pos4), d_ori3, reinterpret_cast<float4 *>(d_pos_gb4),
                     d_image3, d_vel6, d_mask1 ? d_mask1 : ctx->d_mask1, d_noise3, dt, shear_rate, ctx->p.sqm_B1));
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// The second-order Runge-Kutta stages (Integrator.cu:232-438; never called by the reference's step)
int fsd_rk2_stage1(fsd_ctx *ctx, const float *d_pos_in4, const float *d_ori_in3, float *d_pos_out4, float *d_ori_out3,
                   float *d_pos_gb4, const float *d_vel6, const float *d_mask1, const float *d_noise3, float dt,
                   float shear_rate)
{
    RC(k_rk2_stage1(ctx, reinterpret_cast<const float4 *>(d_pos_in4), d_ori_in3, reinterpret_cast<float4 *>(d_pos_out4),
                    d_ori_out3, reinterpret_cast<float4 *>(d_pos_gb4), d_vel6, d_mask1 ? d_mask1 : ctx->d_mask1,
                    d_noise3, dt, shear_rate));
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int fsd_rk2_stage2(fsd_ctx *ctx, float coef1, const float *d_pos_in1_4, const float *d_ori_in1_3, float coef2,
                   const float *d_pos_in2_4, const float *d_ori_in2_3, float *d_pos_out4, float *d_ori_out3,
                   float *d_pos_gb4, const float *d_vel6, const float *d_mask1, const float *d_noise3,
                   int32_t *d_image3, float coef3, float dt, float shear_rate)
{
    RC(k_rk2_stage2(ctx, coef1, reinterpret_cast<const float4 *>(d_pos_in1_4), d_ori_in1_3, coef2,
                    reinterpret_cast<const float4 *>(d_pos_in2_4), d_ori_in2_3, reinterpret_cast<float4 *>(d_pos_out4),
                    d_ori_out3, reinterpret_cast<float4 *>(d_pos_gb4), d_vel6, d_mask1 ? d_mask1 : ctx->d_mask1,
                    d_noise3, d_image3, coef3, dt, shear_rate));
    CK(hipStreamSynchronize(ctx->stream));
    return 0;
}

// initial guess of a step's solve in d_sol: the previous step's solution (Integrator.cu:777), or
// its linear (guess 1) or quadratic (guess 2) extrapolation, of the highest order the solutions
// kept allow.  d_sol = x_{n-1}, d_sol_prev = x_{n-2}, d_sol_prev2 = x_{n-3}
static int initial_guess(fsd_ctx *ctx)
{
    const int64_t n17 = 17 * (int64_t)ctx->N;
    if (ctx->p.guess < 1) return 0;
    {
        const int order = std::min((int)ctx->p.guess, (int)ctx->n_solved - 1);
        if (order >= 1) RC(k_copy(ctx, n17, ctx->d_sol, ctx->d_t));
        if (order == 1) {
            RC(k_axpby(ctx, n17, 2.0f, ctx->d_sol, -1.0f, ctx->d_sol_prev, ctx->d_sol));
        } else if (order == 2) {
            RC(k_axpby(ctx, n17, 3.0f, ctx->d_sol, -3.0f, ctx->d_sol_prev, ctx->d_sol));
            RC(k_axpby(ctx, n17, 1.0f, ctx->d_sol, 1.0f, ctx->d_sol_prev2, ctx->d_sol));
        }
        if (ctx->p.guess >= 2) std::swap(ctx->d_sol_prev, ctx->d_sol_prev2);   // x_{n-2} becomes the older one
        RC(k_copy(ctx, n17, order >= 1 ? ctx->d_t : ctx->d_sol, ctx->d_sol_prev));
    }
    return 0;
}

// Stokes::integrateStepOne -> Stokes_StepOne (Stokes.cc:1197-1683, Stokes.cu:119-273)
int fsd_step(fsd_ctx *ctx, uint32_t timestep, float shear_rate, float *d_pos4, float *d_ori3, float *d_pos_gb4,
             int32_t *d_image3, const float *d_force6, const float *d_noise_ang3, float *d_vel11,
             fsd_solve_stats *stats)
{
    const int N = ctx->N;
    const int64_t n17 = 17 * (int64_t)N;
    float4 *pos = reinterpret_cast<float4 *>(d_pos4);
    if (!ctx->have_table) { ctx->err = "resistance table not set"; return FSD_E_STATE; }
    // neighbour list (HOOMD NeighborList::compute: rebuilt when some particle moved r_buff / 2)
    // and sheared k vectors
    bool rebuilt = false;
    RC(k_update_nlist(ctx, pos, &rebuilt));
    ctx->nlist_builds += rebuilt;
    RC(k_gridk(ctx));
    if (ctx->p.autotune && !ctx->tune.tuned) {   // launch shapes timed at this step's positions
        RC(k_build_lub_pairs(ctx, pos));
        ctx->spread_valid = false;
        RC(autotune_run(ctx, pos));
        ctx->spread_valid = false;
    }
    // preconditioner (Precondition_Wrap): its host phase runs on a second thread while the
    // kernels that do not depend on it are enqueued
    RC(precond_fetch(ctx, pos));
    // precond mode 2: IC(0) calibrates on every lp_calib_every-th step.  The other steps keep the
    // last computed ordering of the particles (fsd_learned_set_keep_order), so that their host plan
    // skips the RCM, while the neighbour list it was computed under stands: no particle has moved
    // r_buff / 2 since, so the pruned graph differs from the one it orders by a few pairs at most
    const bool calib = ctx->precond_mode == 2 && lp_have_weights(ctx) && ctx->lp_calib_every > 0 &&
                       ctx->lp_steps % ctx->lp_calib_every == 0;
    const bool keep_order = ctx->precond_mode == 2 && lp_have_weights(ctx) && ctx->lp_keep_order && !calib &&
                            ctx->order_nlist == ctx->nlist_builds;
    std::thread planner(precond_plan, ctx, keep_order);
    const bool brownian = ctx->p.T > 0.f;
    bool r0_in_t = false;   // the first residual formed below (d_t = A x0 - b, or the refinement's FP64 one)
    auto independent = [&]() -> int {
        // applied force: injected, or the reference force law with the external torque
        if (d_force6) {
            RC(k_copy(ctx, 6 * (int64_t)N, d_force6, ctx->d_force));
        } else {
            const float T_ext = (float)(ctx->p.T_ext * std::cos(ctx->p.omega_ext * (double)timestep * ctx->p.dt));
            RC(k_force_law(ctx, pos, T_ext, ctx->d_force));
        }
        // right-hand side (Integrator_ComputeVelocity, Integrator.cu:744-860)
        RC(k_fill(ctx, 11 * (int64_t)N, 0.f, d_vel11));
        RC(k_fill(ctx, n17, 0.f, ctx->d_rhs));
        RC(k_build_lub_pairs(ctx, pos));
        RC(k_setup_spread(ctx, pos));
        if (!brownian) {
            // without Brownian terms nothing else enters the right-hand side, and the first residual
            // of a right-preconditioned solve needs no preconditioner: both are formed here
            RC(k_strain_rhs(ctx, ctx->d_rhs, shear_rate, d_ori3, ctx->d_force));
            RC(k_rfe(ctx, pos, shear_rate, d_ori3, ctx->d_rhs + 11 * (int64_t)N));
            RC(initial_guess(ctx));
            if (ctx->p.gmres_true_tol > 0.f && !ctx->p.refine) {
                RC(k_saddle_multiply(ctx, pos, ctx->d_sol, ctx->d_t));
                RC(k_axpby(ctx, n17, 1.0f, ctx->d_t, -1.0f, ctx->d_rhs, ctx->d_t));
                r0_in_t = true;
            } else if (ctx->p.gmres_true_tol > 0.f) {   // the refined solve's first FP64 residual
                RC(refine_first_residual(ctx, pos, ctx->d_rhs, ctx->d_sol));
                r0_in_t = true;
            }
        }
        return 0;
    };
    const int rci = independent();
    planner.join();
    if (rci < 0) return rci;
    ctx->lp_force_ichol = calib;
    ctx->lp_skip = ctx->precond_mode == 2 && !ctx->lp_force_ichol && ctx->lp_skip_after > 0 &&
                   ctx->lp_consec_fb >= ctx->lp_skip_after;
    RC(precond_finish(ctx, pos));
    if (brownian) {
        // rhs[0:11N] = U_B^ff, rhs[11N:17N] = F_B^nf, Velocity[0:6N] = RFD drift
        RC(alloc_brownian(ctx));
        RC(bro_farfield(ctx, pos, timestep, ctx->d_rhs));
        RC(bro_nearfield(ctx, pos, timestep, ctx->d_rhs + 11 * (int64_t)N));
        RC(bro_rfd(ctx, pos, timestep, d_vel11));
        RC(k_strain_rhs(ctx, ctx->d_rhs, shear_rate, d_ori3, ctx->d_force));
        RC(k_rfe(ctx, pos, shear_rate, d_ori3, ctx->d_rhs + 11 * (int64_t)N));
        RC(initial_guess(ctx));
    }
    // saddle point solve, warm-started from the previous step's solution
    fsd_solve_stats st;
    memset(&st, 0, sizeof st);
    st.true_residual = -1.f;
    int rc = solve_saddle(ctx, pos, ctx->d_rhs, ctx->d_sol, ctx->p.gmres_maxiter, &st, r0_in_t);
    ctx->n_solved++;
    if (rc < 0) return rc;
    if (ctx->precond_mode == 2) {
        if (ctx->lp_force_ichol) {   // calibration step: solve_saddle recorded IC(0)'s iteration counts
            ctx->lp_calibrations++;
            ctx->lp_consec_fb = 0;   // and the next step tries the learned factor again
        } else if (st.fallback == 2) {
            ctx->lp_consec_fb++;
        } else if (st.precond_used == 1) {
            ctx->lp_consec_fb = 0;
        }
        ctx->lp_force_ichol = false;
        ctx->lp_skip = false;
        ctx->lp_steps++;
    }
    RC(k_axpby(ctx, 6 * (int64_t)N, 1.0f, ctx->d_sol + 11 * (int64_t)N, 1.0f, d_vel11, d_vel11));
    if (ctx->p.period > 0 && (int)(timestep + 1) % ctx->p.period == 0) {
        float *S = d_vel11 + 6 * (int64_t)N;
        RC(k_axpby(ctx, 5 * (int64_t)N, 1.0f, ctx->d_sol + 6 * (int64_t)N, 1.0f, S, S));
        RC(k_rsu(ctx, pos, d_vel11, S));
        RC(k_rse(ctx, pos, shear_rate, d_ori3, S));
    }
    // explicit Euler with shear (Integrator_ExplicitEuler_Shear_kernel)
    RC(k_euler_shear(ctx, pos, d_ori3, reinterpret_cast<float4 *>(d_pos_gb4), d_image3, d_vel11, ctx->d_mask1,
                     d_noise_ang3, ctx->p.dt, shear_rate, ctx->p.sqm_B1));
    ctx->lub_valid = false;      // the pair records and the spreading set-up describe the old positions
    ctx->spread_valid = false;
    CK(hipStreamSynchronize(ctx->stream));
    st.ichol_relaxer = ctx->ichol_relaxer;
    st.ichol_attempts = ctx->ichol_attempts;
    st.nnz = ctx->pch.nnz_ref;
    st.nlist_total = ctx->nlist_total;
    st.lub_pairs = ctx->lub_total;
    st.lower_blocks = ctx->pch.nlow;
    st.levels = ctx->pch.nlev_f;
    st.sweep_steps = ctx->pch.tp_ok ? ctx->pch.tp_f.nsteps + ctx->pch.tp_b.nsteps : 0;
    st.nlist_rebuilt = rebuilt ? 1 : 0;
    st.lanczos_ff = (ctx->p.T > 0.f) ? ctx->last_lanczos_ff : 0;
    st.lanczos_nf = (ctx->p.T > 0.f) ? ctx->last_lanczos_nf : 0;
    // relaxer and Krylov-size bookkeeping (Stokes.cc:1658-1666)
    ctx->ichol_relaxer = ((timestep % 5 == 0) || (ctx->ichol_relaxer > 1024.0f)) ? 1.0f : ctx->ichol_relaxer;
    if (ctx->p.T > 0.f) {
        ctx->m_lanczos_ff = ((timestep % 100 == 0) || (ctx->m_lanczos_ff > 50)) ? 3 : ctx->m_lanczos_ff;
        ctx->m_lanczos_nf = ((timestep % 100 == 0) || (ctx->m_lanczos_nf > 50)) ? 3 : ctx->m_lanczos_nf;
    }
    if (stats) *stats = st;
    return rc;
}

int fsd_profile_enable(fsd_ctx *ctx, int32_t enable)
{
    ctx->prof = (unsigned)enable;
    return 0;
}

int fsd_profile_every(fsd_ctx *ctx, int32_t every)
{
    if (every < 1) { ctx->err = "fsd_profile_every: every must be at least 1"; return FSD_E_PARAM; }
    ctx->prof_every = every;
    for (auto &s : ctx->slots) s.seen = 0;
    return 0;
}

int fsd_profile_get(fsd_ctx *ctx, const char *name, int64_t *launches, double *total_ms)
{
    prof_collect(ctx);
    for (auto &s : ctx->slots)
        if (strcmp(s.name, name) == 0) {
            *launches = s.launches;
            *total_ms = s.ms;
            return 0;
        }
    return FSD_E_PARAM;
}

int fsd_profile_reset(fsd_ctx *ctx)
{
    prof_collect(ctx);
    for (auto &s : ctx->slots) { s.launches = 0; s.ms = 0.0; }
    return 0;
}

}  // extern "C"

// ------------------------------------------------------ learned factor (mode 2)
int fsd_learned_nparams(fsd_ctx *ctx, int64_t *n)
{
    RC(lp_init(ctx));
    *n = lp_nparam(ctx);
    return 0;
}

int fsd_learned_set_weights(fsd_ctx *ctx, const float *h_w, int64_t n) { return lp_set_weights(ctx, h_w, n); }

int fsd_learned_get_weights(fsd_ctx *ctx, float *h_w, int64_t n) { return lp_get_weights(ctx, h_w, n, false); }

int fsd_learned_get_grad(fsd_ctx *ctx, float *h_g, int64_t n) { return lp_get_weights(ctx, h_g, n, true); }

int fsd_learned_get_adam(fsd_ctx *ctx, float *h_m, float *h_v, int64_t n)
{
    return lp_adam_state(ctx, h_m, h_v, n, false);
}

int fsd_learned_set_adam(fsd_ctx *ctx, const float *h_m, const float *h_v, int64_t n)
{
    return lp_adam_state(ctx, const_cast<float *>(h_m), const_cast<float *>(h_v), n, true);
}

int fsd_learned_set_calibration(fsd_ctx *ctx, int32_t every)
{
    if (every < 0) { ctx->err = "calibration period must be >= 0"; return FSD_E_PARAM; }
    ctx->lp_calib_every = every;
    ctx->lp_steps = 0;
    return 0;
}

int fsd_learned_set_keep_order(fsd_ctx *ctx, int32_t on)
{
    ctx->lp_keep_order = on != 0;
    return 0;
}

int fsd_learned_set_skip(fsd_ctx *ctx, int32_t after)
{
    if (after < 0) { ctx->err = "skip count must be >= 0"; return FSD_E_PARAM; }
    ctx->lp_skip_after = after;
    ctx->lp_consec_fb = 0;
    return 0;
}

int fsd_learned_inject_pivot(fsd_ctx *ctx, int32_t particle)
{
    if (particle < 0 || particle >= ctx->N) { ctx->err = "particle out of range"; return FSD_E_PARAM; }
    ctx->lp_inject = particle;
    return 0;
}

int fsd_learned_stats(fsd_ctx *ctx, int64_t *h_out6)
{
    h_out6[0] = ctx->lp_steps;
    h_out6[1] = ctx->lp_pivot_fallbacks;
    h_out6[2] = ctx->lp_iter_fallbacks;
    h_out6[3] = ctx->lp_calibrations;
    float lambda = 1.0f;
    int shrunk = 0, plain = 0;
    lp_last_factor(ctx, &lambda, &shrunk, &plain);
    h_out6[4] = shrunk;
    h_out6[5] = plain;
    return 0;
}

int fsd_learned_loss_grad(fsd_ctx *ctx, const float *d_pos4, int32_t nsamp, const float *d_x17, const float *d_b17,
                          int32_t want_grad, double *loss)
{
    if (ctx->precond_mode != 2) { ctx->err = "the learned factor needs precond mode 2"; return FSD_E_STATE; }
    const float4 *pos = reinterpret_cast<const float4 *>(d_pos4);
    RC(current_nlist(ctx, pos));
    RC(precond_fetch(ctx, pos));
    precond_plan(ctx);
    RC(precond_upload_blocks(ctx, pos));
    prof_end(ctx, PROF_PRECOND_BUILD, ctx->pc_ev);
    int failed = 0;
    RC(lp_forward(ctx, pos, true, -1, &failed));
    lp_set_active(ctx, !failed);
    ctx->precond_valid = !failed;
    if (failed) {
        ctx->err = "learned factor has a non-positive pivot";
        return FSD_W_PIVOT;
    }
    RC(lp_loss_grad(ctx, nsamp, d_x17, d_b17, loss, want_grad != 0));
    return 0;
}

int fsd_learned_export(fsd_ctx *ctx, float *h_oute, float *h_outn, float *h_W, float *h_Dinv, int64_t edge_cap)
{
    if (!lp_active(ctx)) { ctx->err = "no learned factor is active"; return FSD_E_STATE; }
    if (ctx->pch.nlow > edge_cap) return FSD_E_CAPACITY;
    return lp_export(ctx, h_oute, h_outn, h_W, h_Dinv);
}

int fsd_learned_levels(fsd_ctx *ctx, int32_t *h_levels, int64_t n)
{
    if (!ctx->lp) { ctx->err = "no learned factor"; return FSD_E_STATE; }
    return lp_levels(ctx, h_levels, n);
}

int fsd_learned_factor_loss(fsd_ctx *ctx, const float *d_pos4, int32_t want_grad, double *loss)
{
    if (ctx->precond_mode != 2) { ctx->err = "the learned factor needs precond mode 2"; return FSD_E_STATE; }
    const float4 *pos = reinterpret_cast<const float4 *>(d_pos4);
    RC(current_nlist(ctx, pos));
    RC(precond_fetch(ctx, pos));
    precond_plan(ctx);
    RC(precond_upload_blocks(ctx, pos));
    prof_end(ctx, PROF_PRECOND_BUILD, ctx->pc_ev);
    // the targets: IC(0) with the relaxer doubling from 1, on this pattern
    ctx->ichol_attempts = 0;
    float lambda = 1.0f;
    for (;;) {
        int failed = 0;
        RC(k_ichol(ctx, lambda, &failed));
        ctx->ichol_attempts++;
        if (!failed) break;
        lambda *= 2.0f;
        if (lambda > 1.0e7f) { ctx->err = "IC(0) did not succeed for any relaxer"; return FSD_E_ICHOL; }
    }
    int failed = 0;
    RC(lp_forward(ctx, pos, true, -1, &failed));
    lp_set_active(ctx, false);
    ctx->precond_valid = false;
    if (failed) {
        ctx->err = "learned factor has a non-positive pivot";
        return FSD_W_PIVOT;
    }
    RC(lp_factor_loss_grad(ctx, loss, want_grad != 0));
    return 0;
}

int fsd_learned_revert(fsd_ctx *ctx)
{
    if (!ctx->lp) { ctx->err = "no learned factor"; return FSD_E_STATE; }
    return lp_revert(ctx);
}

int fsd_learned_adam(fsd_ctx *ctx, float lr, int32_t step)
{
    if (step < 1) { ctx->err = "Adam step counts from 1"; return FSD_E_PARAM; }
    RC(lp_init(ctx));
    return lp_adam(ctx, lr, step);
}

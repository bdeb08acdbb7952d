// Internal declarations of the MI355X FSD library (libfsd.so).
//
// Data layout in HBM (all device arrays owned by the context unless noted):
//   particles   pos float4[N], generalised vectors in the reference layouts;
//   neighbours  fixed-stride list nlist[i*maxn + k] with counts nneigh[i];
//   lubrication per-pair records (neighbour, unit vector, five "12" resistance
//               functions) + one symmetric 6x6 self block per particle,
//               built once per step and reused by every R_FU apply;
//   PSE grids   11 real fields [11][M] (M = Nx*Ny*Nz, z fastest) and their
//               half spectra [11][Nx*Ny*(Nz/2+1)] (rocFFT R2C/C2R);
//   IC(0)       block-sparse lower factor on the RCM-permuted particle graph:
//               per permuted particle a dense 6x6 diagonal block and one 6x6
//               block per lower neighbour (36 floats, row-major).
#pragma once
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "fsd.h"
#include "precond_plan.h"

#define FSD_MAXP 14        // pruned neighbours incl. self (kissing number 12 + self + 1)
#define FSD_PSTRIDE 16     // stride of the pruned list
#define BRO_MMAX 100       // Lanczos basis bound of the Brownian samplers (m_max, Brownian_FarField.cu:613)

struct BoxF {
    float Lx, Ly, Lz, xy, xz, yz;
    float iLx, iLy, iLz;
};

// HOOMD BoxDim::minImage with rint rounding, evaluated without contraction so
// that cutoff decisions round exactly like the oracle.
__host__ __device__ inline void box_min_image(const BoxF &b, float &dx, float &dy, float &dz)
{
#pragma clang fp contract(off)
    float img = rintf(dz * b.iLz);
    float t = b.Lz * b.xz; t = t * img; dx = dx - t;
    t = b.Lz * b.yz; t = t * img; dy = dy - t;
    t = b.Lz * img; dz = dz - t;
    img = rintf(dy * b.iLy);
    t = b.Ly * b.xy; t = t * img; dx = dx - t;
    t = b.Ly * img; dy = dy - t;
    img = rintf(dx * b.iLx);
    t = b.Lx * img; dx = dx - t;
}

__host__ __device__ inline float dist2_nc(float dx, float dy, float dz)
{
#pragma clang fp contract(off)
    float s = dx * dx;
    float t = dy * dy;
    s = s + t;
    t = dz * dz;
    return s + t;
}

// HOOMD BoxDim::makeFraction
__host__ __device__ inline void box_make_fraction(const BoxF &b, float px, float py, float pz,
                                                  float &fx, float &fy, float &fz)
{
#pragma clang fp contract(off)
    float dx = px + 0.5f * b.Lx, dy = py + 0.5f * b.Ly, dz = pz + 0.5f * b.Lz;
    float t = b.yz * b.xy; t = b.xz - t; t = t * pz;
    float u = b.xy * py; t = t + u; dx = dx - t;
    u = b.yz * pz; dy = dy - u;
    fx = dx / b.Lx; fy = dy / b.Ly; fz = dz / b.Lz;
}

// HOOMD BoxDim::wrap (one image shift per axis, x then y then z)
__host__ __device__ inline void box_wrap(const BoxF &b, float &x, float &y, float &z, int &ix, int &iy, int &iz)
{
    const float hx = 0.5f * b.Lx, hy = 0.5f * b.Ly, hz = 0.5f * b.Lz;
    float tilt_x = (b.xz - b.xy * b.yz) * z + b.xy * y;
    if (x >= hx + tilt_x) { x -= b.Lx; ix++; }
    else if (x < -hx + tilt_x) { x += b.Lx; ix--; }
    float tilt_y = b.yz * z;
    if (y >= hy + tilt_y) { y -= b.Ly; x -= b.Ly * b.xy; iy++; }
    else if (y < -hy + tilt_y) { y += b.Ly; x += b.Ly * b.xy; iy--; }
    if (z >= hz) { z -= b.Lz; y -= b.Lz * b.yz; x -= b.Lz * b.xz; iz++; }
    else if (z < -hz) { z += b.Lz; y += b.Lz * b.yz; x += b.Lz * b.xz; iz--; }
}

// Mobility_SetGridk_kernel (Helper_Mobility.cu:351-410) for the full-grid
// index (i, j, k): sheared wave vector and the Hasimoto factor
// w = 6 pi (1 + k^2/4xi^2) exp(-(1 - eta) k^2/4xi^2) / k^2 / (Nx Ny Nz)
// evaluated in double, stored as float (w = 0 at the origin)
__host__ __device__ inline float4 gridk_value(int i, int j, int k, int Nx, int Ny, int Nz, const BoxF &b, float xi,
                                              float eta, float twopi_k)
{
    double kx = (i < (Nx + 1) / 2) ? i : i - Nx;
    double ky = (((j < (Ny + 1) / 2) ? j : j - Ny) - (double)b.xy * kx * b.Ly / b.Lx) / b.Ly;
    double kz = ((k < (Nz + 1) / 2) ? k : k - Nz) / (double)b.Lz;
    kx = kx / b.Lx;
    kx *= twopi_k; ky *= twopi_k; kz *= twopi_k;
    const double k2 = kx * kx + ky * ky + kz * kz;
    const double xisq = (double)xi * xi;
    double w = 0.0;
    if (!(i == 0 && j == 0 && k == 0))
        w = 6.0 * 3.1415926536 * (1.0 + k2 / 4.0 / xisq) * exp(-(1.0 - eta) * k2 / 4.0 / xisq) / k2
            / (double)((int64_t)Nx * Ny * Nz);
    return make_float4((float)kx, (float)ky, (float)kz, (float)w);
}

// True when the mode (i, j, k) has a Nyquist index: there the full grid's
// conjugate mode carries the same Nyquist k component (not its negative), so
// a scaled spectrum is not Hermitian and the real part of the reference's C2C
// inverse equals the inverse of the Hermitian part (X(k) + conj X(k*)) / 2.
__host__ __device__ inline bool has_nyquist(int i, int j, int k, int Nx, int Ny, int Nz)
{
    return (Nx % 2 == 0 && i == Nx / 2) || (Ny % 2 == 0 && j == Ny / 2) || (Nz % 2 == 0 && k == Nz / 2);
}

// Workgroups are dispatched to the 8 XCDs of an MI355X round-robin (workgroup
// w runs on XCD w mod 8), and each XCD has its own L2.  xcd_remap maps
// workgroup w of a grid of n to a work item so that every XCD gets one
// contiguous range of items: neighbouring items (which share halo data)
// then share an L2.  It is a bijection of [0, n) whatever the dispatch.
constexpr int FSD_NXCD = 8;
__device__ inline int xcd_remap(int w, int n)
{
    const int q = n / FSD_NXCD, r = n % FSD_NXCD;
    const int x = w % FSD_NXCD, k = w / FSD_NXCD;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

// Sum over aligned groups of G lanes (G = 2, 4, 8, 16) with DPP lane
// exchanges inside a 16-lane row (quad_perm, row_half_mirror, row_mirror):
// VALU only, no LDS round trip.  Every lane of a group gets the group's sum.
template <int CTRL>
__device__ __forceinline__ float dpp_get(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}
template <int G>
__device__ __forceinline__ float group_sum(float v)
{
    static_assert(G == 2 || G == 4 || G == 8 || G == 16, "group of 2, 4, 8 or 16 lanes");
    v += dpp_get<0xB1>(v);                       // quad_perm(1, 0, 3, 2): lane ^ 1
    if constexpr (G >= 4) v += dpp_get<0x4E>(v);  // quad_perm(2, 3, 0, 1): lane ^ 2
    if constexpr (G >= 8) v += dpp_get<0x141>(v); // row_half_mirror: the other quad of the 8
    if constexpr (G >= 16) v += dpp_get<0x140>(v); // row_mirror: the other 8 of the 16
    return v;
}

struct LubTable {
    const float *dist;   // 1000
    const float *vals;   // 22000
    float table_min, table_dr;
};

// Interpolate `ncol` resistance functions starting at column c0
// (Lubrication.cu:172-231); contact_row < 0 disables the r <= 2.001 clamp.
__device__ inline void lub_lookup(const LubTable &t, float dist, int c0, int ncol, int contact_row, float *out)
{
    if ((double)dist <= 2.001 && contact_row >= 0) {
        for (int c = 0; c < ncol; ++c) out[c] = t.vals[22 * contact_row + c0 + c];
        return;
    }
    float arg = (float)(((double)dist - 2.0) / (double)t.table_min);
    int ind = (int)(log10f(arg) / t.table_dr);
    ind = ind < 0 ? 0 : ind;
    if (ind >= 999) {  // r in [~3.99968, 4): the reference reads past the distance array; row 999 is zero
        for (int c = 0; c < ncol; ++c) out[c] = t.vals[22 * 999 + c0 + c];
        return;
    }
    float dl = t.dist[ind], du = t.dist[ind + 1];
    float fac = (dist - dl) / (du - dl);
    for (int c = 0; c < ncol; ++c) {
        float lo = t.vals[22 * ind + c0 + c], hi = t.vals[22 * (ind + 1) + c0 + c];
        out[c] = lo + (hi - lo) * fac;
    }
}

// ------------------------------------------------------------------ context
struct ProfSlot {
    const char *name;
    int64_t launches;             // launches timed
    int64_t seen = 0;             // launches of the class while enabled (one in prof_every is timed)
    double ms;
    struct Pending { hipEvent_t b, e; int n; };   // event pair around n launches of the class
    std::vector<Pending> pending;
};

// Launch shapes chosen by the autotuner (fsd_autotune): each knob selects one of the
// kernels or launch shapes that compute the same operator.
struct fsd_tuning {
    int spread = 0;          // 0: spread_tile2_kernel (tile 8, P 6), 1: spread_kernel (one workgroup per tile)
    int contract = 0;        // 0: tile contraction (contract_tile2/contract_tile, when it fits in LDS),
                             // 1: one wave per particle (contract_kernel)
    int xpass = 0;           // 0: x transforms fused with the Green's function (when built), 1: 3-D rocFFT
    int real_g = 16, real_blk = 256;   // real-space sum: lanes per particle (4, 8, 16), workgroup size
    int rfu_g = 8, rfu_blk = 256;      // R_FU product: lanes per particle (8, 16), workgroup size
    int tuned = 0;
    float us[3] = {0.f, 0.f, 0.f};     // best time per apply (us): wave part, real part, R_FU
};

struct fsd_ctx {
    fsd_params p;
    fsd_pse_info pse;
    fsd_tuning tune;
    std::string err;
    hipStream_t stream = nullptr;
    BoxF box;
    int N = 0;

    // tables
    float *d_ewald = nullptr; int ewald_nR = 0;
    std::vector<float> h_ewald;
    float *d_res_dist = nullptr, *d_res_vals = nullptr; bool have_table = false;
    std::vector<float> h_res_dist, h_res_vals;   // host copy (friction modifies the values)
    float table_min = 0.0001f, table_dr = 0.004305f;

    // neighbour list
    int maxn = 0;
    int *d_nneigh = nullptr, *d_nlist = nullptr;
    int ncell[3] = {0, 0, 0}; int ncells = 0; bool use_cells = false;
    int *d_cell_of = nullptr, *d_cell_count = nullptr, *d_cell_start = nullptr, *d_cell_list = nullptr;
    int *d_flags = nullptr;     // [0] nlist overflow, [1] lub overflow, [2] prune overflow, [4] nlist entries,
                                // [5] lubrication pairs, [6] some particle moved r_buff / 2 since the list
    float4 *d_pos_last = nullptr;   // positions the neighbour list was built at
    int64_t nlist_builds = 0;       // neighbour lists built by this context
    int *h_flags = nullptr;     // pinned mirror
    bool nlist_valid = false;
    int64_t nlist_total = 0;

    // lubrication pair records
    int maxlub = 0;
    int *d_lub_n = nullptr, *d_lub_j = nullptr;
    float4 *d_lub_a = nullptr;  // (rx, ry, rz, XA12-YA12)
    float4 *d_lub_b = nullptr;  // (YA12, YB12, XC12-YC12, YC12)
    float *d_lub_self = nullptr; // 36 per particle (row-major 6x6)
    bool lub_valid = false;
    int64_t lub_total = 0;

    // PSE
    int M = 0, Mh = 0;
    float *d_grid = nullptr;      // 11*M real
    float2 *d_spec = nullptr;     // 11*Mh complex
    float4 *d_gridk = nullptr;    // Mh
    rocfft_plan plan_fwd = nullptr, plan_inv = nullptr;
    // 2-D (y, z) transforms batched over the fields and x planes; the x direction and the Green's
    // function then run in one kernel (fx_B: Nx / 8, 0 when the 3-D plans are used instead)
    rocfft_plan plan_fwd2 = nullptr, plan_inv2 = nullptr;
    int fx_B = 0;
    rocfft_execution_info info_fwd = nullptr, info_inv = nullptr;
    void *d_fftwork = nullptr; size_t fftwork_size = 0;
    int tiles[3] = {0, 0, 0}; int ntiles = 0; int tile = 8;
    int *d_porig = nullptr;       // 3 ints per particle: stencil origin (unwrapped)
    int *d_ocell = nullptr;       // N: grid cell of each particle's (wrapped) stencil origin
    int *d_olist = nullptr;       // N: particles sorted by origin cell, ascending id inside a cell
    int *d_ocount = nullptr;      // M: particles per origin cell (scratch)
    int *d_ostart = nullptr;      // M+1: first position of each origin cell in d_olist
    int *d_scan_tmp = nullptr;    // chunk totals of the multi-workgroup scan
    int contract_lds = 0;         // LDS bytes of the tile contraction (0: per-particle contraction)
    bool contract_split = false;  // the two-workgroup-per-tile contraction (P = 6, tile 8)
    float *d_wave = nullptr;      // 11 per particle: wave-space u (3) and gradient (8)
    float4 *d_rec = nullptr;      // 3 per particle: force + couplet records of the current apply
    bool spread_valid = false;
    const float *spread_pos = nullptr;

    // preconditioner (blocks on the permuted particle graph)
    int *d_pn = nullptr, *d_plist = nullptr, *d_pless = nullptr;
    int *d_perm = nullptr, *d_iperm = nullptr;
    std::vector<int> h_pn;
    int *d_lrow = nullptr;        // N+1: start of each permuted particle's lower blocks
    int *d_lcol = nullptr;        // lower block columns (permuted ids), ascending
    int *d_urow = nullptr, *d_ucol = nullptr, *d_upos = nullptr;  // upper adjacency for L^T solves
    float *d_Ablk = nullptr;      // 36 per lower block (A values)
    float *d_Adiag = nullptr;     // 36 per particle (A diagonal block, full)
    float *d_Lblk = nullptr, *d_Ldiag = nullptr;
    int *d_isolated = nullptr;    // per permuted particle
    int64_t nlow_cap = 0;         // lower blocks allocated
    float ichol_relaxer = 1.0f;
    int ichol_attempts = 0;
    bool precond_valid = false;
    int *d_lev_f = nullptr, *d_lev_b = nullptr;
    float *d_tmp6a = nullptr, *d_tmp6b = nullptr;
    float *d_g12 = nullptr;       // G b_F and G b_U of the saddle preconditioner
    // host phase of the preconditioner build (kept between steps to reuse capacity)
    std::vector<int> h_plstride;                         // pruned lists, FSD_PSTRIDE per particle
    PrecondHost pch;                                     // RCM, pattern, levels, sweep plans
    hipEvent_t pc_ev = nullptr;                          // profiling scope of the build
    // single-workgroup sweeps: task plans (sweep_plan.h) and their device operands
    int tp_R = 0;                                        // LDS ring capacity in rows (0: level path)
    bool tp_joint = false;                               // both right-hand sides in one workgroup
    int64_t tp_cap_f = 0, tp_cap_b = 0;                  // lanes allocated per sweep
    int2 *d_tp_finfo = nullptr, *d_tp_binfo = nullptr;   // per lane
    int2 *d_tp_fbsrc = nullptr, *d_tp_bbsrc = nullptr;   // per lane
    float4 *d_tp_fblk = nullptr, *d_tp_bblk = nullptr;   // 9 per lane
    float4 *d_tp_finp = nullptr, *d_tp_binp = nullptr;   // 3 N: per-row inputs
    float4 *d_tp_fres = nullptr, *d_tp_bres = nullptr;   // 3 N: per-row results
    int *d_tp_fseq = nullptr, *d_tp_bseq = nullptr;      // N: sequence number of each permuted particle
    int *d_tp_fplace = nullptr, *d_tp_bplace = nullptr;  // 3 N: first lane and step of each row, step ends
    float *d_tp_Z = nullptr;                             // 36 N: inverse diagonal blocks
    // application by Jacobi iterations (precond modes 1 and 2)
    int precond_mode = 0;                                // 0: exact sweeps, 1: IC(0) + Jacobi, 2: learned + Jacobi
    int jac_k = 0;                                       // Jacobi iterations per triangular solve
    int ic_jac_k = 0;                                    // the same for IC(0) (modes 1 and 2); 0: exact solves
    float *d_Fblk = nullptr;                             // 36 per lower block: Z_P L_PQ
    int *d_lrowidx = nullptr;                            // row P of every lower block
    float *d_jac = nullptr;                              // 4 x 6N iterates
    float *d_Lblk2 = nullptr, *d_Ldiag2 = nullptr;       // second factor buffer of the parallel IC(0) sweeps
    // learned factor (precond mode 2, learned.h) and its fallbacks to IC(0)
    struct LearnedState *lp = nullptr;
    int lp_calib_every = 50;       // every this many steps one step runs IC(0) to measure its iteration count
    int lp_iter_cap[3] = {0, 0, 0};   // IC(0)'s iterations at the last such step (0: none yet), for the
                                      // step's solve and the two random-finite-difference solves
    int lp_inject = -1;            // particle whose learned pivot the next factor forces non-positive (tests)
    bool lp_force_ichol = false;   // this step's preconditioner is IC(0) (calibration)
    bool lp_keep_order = true;     // steps keep the last ordering while its neighbour list stands
    int64_t order_nlist = -1;      // nlist_builds when the preconditioner's ordering was last computed
    int lp_skip_after = 2;         // this many iteration fallbacks in a row skip the learned factor until the
                                   // next calibration step (0: never skip)
    int lp_consec_fb = 0;          // iteration fallbacks in a row since the last learned solve that converged
    bool lp_skip = false;          // this step skips the learned factor (IC(0), fallback 3)
    int lp_step_fallback = 0;      // 1: this step's learned factor had a non-positive pivot, 3: skipped
    int64_t lp_steps = 0, lp_pivot_fallbacks = 0, lp_iter_fallbacks = 0, lp_calibrations = 0;

    // saddle / GMRES
    float *d_V = nullptr, *d_w = nullptr, *d_t = nullptr, *d_t2 = nullptr;
    float *d_Z = nullptr;         // M^-1 V of the right-preconditioned solve (allocated on first use)
    double *d_part = nullptr; int part_blocks = 0; double *h_part = nullptr;
    hipEvent_t gm_ev[2] = {nullptr, nullptr};   // readbacks of the Hessenberg columns (two slots)
    int kmax = 64;                // most vectors of one multidot (restart + 1, at least 64)
    float *d_rhs = nullptr, *d_sol = nullptr;
    // FP64-residual iterative refinement (allocated on first use)
    double *d_x64 = nullptr, *d_r64 = nullptr, *d_ybot64 = nullptr, *d_norm64 = nullptr;
    double ref_r0[2] = {0.0, 0.0};   // |r|^2, |b|^2 of a first residual formed ahead of the solve
    float *d_refx32 = nullptr, *d_refy32 = nullptr, *d_refr32 = nullptr, *d_refd = nullptr;
    float *d_sol_prev = nullptr;  // x_{n-2} when the initial guess extrapolates (guess >= 1)
    float *d_sol_prev2 = nullptr; // x_{n-3} for the quadratic extrapolation (guess = 2)
    int64_t n_solved = 0;         // steps solved by this context
    float *d_force = nullptr;
    float *d_mask1 = nullptr, *d_mask2 = nullptr;
    float *d_ori_zero = nullptr;

    // Brownian motion (allocated on the first step with T > 0)
    bool bro_alloc = false;
    int m_lanczos_ff = 2, m_lanczos_nf = 2;   // Krylov sizes carried between steps (Stokes.cc:122-123)
    float *d_lanV = nullptr;      // (BRO_MMAX + 1) x 11N Lanczos basis
    float *d_lanw = nullptr, *d_lany0 = nullptr;           // 11N each
    double *d_lan_h = nullptr, *h_lan = nullptr;           // 256 doubles (device, pinned host)
    float *d_noise = nullptr;     // generated noise: 11N far real | 6M far grid | 6N near | 6N RFD
    const float *inj_ff_real = nullptr, *inj_ff_grid = nullptr, *inj_nf = nullptr, *inj_rfd = nullptr;
    float4 *d_posp = nullptr;     // shifted positions of the random finite difference
    float *d_rfd_rhs = nullptr, *d_rfd_sol = nullptr;     // 17N each
    int last_lanczos_ff = 0, last_lanczos_nf = 0;

    // profiling
    unsigned prof = 0;            // bit mask of timed kernel classes
    std::vector<ProfSlot> slots;
    int prof_every = 1;
};

// kernel-launch helper with error capture
#define FSD_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);               \
            return FSD_E_HIP;                                                           \
        }                                                                               \
    } while (0)

static inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

// profiling scopes
void prof_begin(fsd_ctx *ctx, int slot, hipEvent_t *ev);
void prof_end(fsd_ctx *ctx, int slot, hipEvent_t ev);
// end of a scope that enclosed n back-to-back launches of the class (one event pair for the
// batch, so that the event overhead is shared by n launches of a short kernel)
void prof_end_n(fsd_ctx *ctx, int slot, hipEvent_t ev, int n);
enum { PROF_SPREAD = 0, PROF_GREEN, PROF_CONTRACT, PROF_FFT, PROF_REAL, PROF_RFU, PROF_TRSV, PROF_ICHOL,
       PROF_GMRES_VEC, PROF_NLIST, PROF_PRECOND_BUILD, PROF_SWEEP, PROF_PLAN_HOST, PROF_FINISH_HOST,
       PROF_LJIT,          // the forward block Jacobi iterations of the learned factor (k per application)
       PROF_NET,           // features, network and factor assembly of the learned factor
       PROF_JAC,           // the forward block Jacobi iterations on the IC(0) factor (k per application)
       PROF_COUNT };
// host wall time of a phase, added to a slot when that slot is enabled
void prof_host(fsd_ctx *ctx, int slot, double ms);

// kernel groups (launchers live with their kernels)
int k_build_nlist(fsd_ctx *ctx, const float4 *pos);
int k_update_nlist(fsd_ctx *ctx, const float4 *pos, bool *rebuilt);
int k_build_lub_pairs(fsd_ctx *ctx, const float4 *pos);
int k_lub_pairs_into(fsd_ctx *ctx, const float4 *pos, int maxlub, int *lub_n, int *lub_j, float4 *lub_a,
                     float4 *lub_b, float *lub_self, int *overflow);
int k_rfu_apply_with(fsd_ctx *ctx, const float *U, float *F, const int *lub_n, const int *lub_j, const float4 *lub_a,
                     const float4 *lub_b, const float *lub_self, int maxlub);
int noalloc_init_device();   // k_noalloc.hip: device address of the static scratch, once per device
int k_rfu_apply(fsd_ctx *ctx, const float *U, float *F, const float *xF, float sign_out);
int k_rfe(fsd_ctx *ctx, const float4 *pos, float shear_rate, const float *ori, float *F);
int k_rsu(fsd_ctx *ctx, const float4 *pos, const float *U, float *S);
int k_rse(fsd_ctx *ctx, const float4 *pos, float shear_rate, const float *ori, float *S);
int k_setup_spread(fsd_ctx *ctx, const float4 *pos);
int k_contract_lds(fsd_ctx *ctx);
int scan_excl_chunks(int n);
int k_gridk(fsd_ctx *ctx);
float k_twopi(const fsd_ctx *ctx);
int k_mobility(fsd_ctx *ctx, const float4 *pos, const float *gF, float *gU, int parts, const float *addU);
int k_prune(fsd_ctx *ctx, const float4 *pos);
int k_build_blocks(fsd_ctx *ctx, const float4 *pos);
int k_ichol(fsd_ctx *ctx, float lambda, int *failed);
int k_ichol_sweeps(fsd_ctx *ctx, float lambda, int sweeps, int *failed);
int k_sweep_setup(fsd_ctx *ctx);
int k_sweep_pack(fsd_ctx *ctx);
int k_normals(fsd_ctx *ctx, int64_t n, uint32_t step, uint32_t stream, float sd, float *out);
int k_ff_noise(fsd_ctx *ctx, const float *noise, float fac);
int k_wave_to_general(fsd_ctx *ctx, float *out);
int k_shift_wrap(fsd_ctx *ctx, const float4 *pos, const float *psi6, float scale, float4 *out);
int k_wave_contract(fsd_ctx *ctx, const float4 *pos, bool inverse_yz = false);
int k_precond_apply2(fsd_ctx *ctx, const float *b0, const float *b1, float *y0, float *y1, int n_rhs);
int k_export_csr(fsd_ctx *ctx, const float4 *pos, int32_t *h_rowptr, int32_t *h_col, float *h_val,
                 int32_t *h_prow, int32_t *h_pcol, float *h_L, int64_t cap);
int k_force_law(fsd_ctx *ctx, const float4 *pos, float T_ext, float *F);
int k_strain_rhs(fsd_ctx *ctx, float *rhs, float shear_rate, const float *ori, const float *force);
int k_rk2_stage1(fsd_ctx *ctx, const float4 *pos_in, const float *ori_in, float4 *pos_out, float *ori_out,
                 float4 *pos_gb, const float *vel, const float *mask1, const float *noise, float dt, float shear_rate);
int k_rk2_stage2(fsd_ctx *ctx, float c1, const float4 *pos1, const float *ori1, float c2, const float4 *pos2,
                 const float *ori2, float4 *pos_out, float *ori_out, float4 *pos_gb, const float *vel,
                 const float *mask1, const float *noise, int *image, float c3, float dt, float shear_rate);
int k_euler_shear(fsd_ctx *ctx, float4 *pos, float *ori, float4 *pos_gb, int *image, const float *vel,
                  const float *mask1, const float *noise, float dt, float shear_rate, float B1);
// vector helpers
int k_axpby(fsd_ctx *ctx, int64_t n, float a, const float *x, float b, const float *y, float *z);
int k_copy(fsd_ctx *ctx, int64_t n, const float *x, float *y);
int k_fill(fsd_ctx *ctx, int64_t n, float v, float *y);
int autotune_run(fsd_ctx *ctx, const float4 *pos);
int k_multidot(fsd_ctx *ctx, int64_t n, int k, const float *V, int64_t ldv, const float *w, double *out);
int k_multi_axpy(fsd_ctx *ctx, int64_t n, int k, const float *V, int64_t ldv, const double *h, float *w);
int k_nrm2(fsd_ctx *ctx, int64_t n, const float *x, double *out);
int k_scale_by_norm(fsd_ctx *ctx, int64_t n, float *v, const double *h, int K, double *hn_out);
int k_saddle_precond(fsd_ctx *ctx, const float *b, float *x);
// the factor's application by Jacobi iterations (k_jacobi.hip)
int k_diag_inverse(fsd_ctx *ctx);
int k_jacobi_pack(fsd_ctx *ctx);
int k_saddle_precond_jacobi(fsd_ctx *ctx, const float *b, float *x);
int k_precond_apply_jacobi(fsd_ctx *ctx, const float *b, float *y);
int k_saddle_multiply(fsd_ctx *ctx, const float4 *pos, const float *x, float *y);
// FP64-residual refinement (k_refine.hip)
int k_residual64(fsd_ctx *ctx, const float4 *pos, const float *b, const double *x64, float *x32, float *y32,
                 double *ybot, double *r64, float *r32, double *norms);
int k_to64(fsd_ctx *ctx, int64_t n, const float *x, double *y);
int k_to32(fsd_ctx *ctx, int64_t n, const double *x, float *y);
int k_add64(fsd_ctx *ctx, int64_t n, const float *d, double *x);
// the learned factor (k_learned.hip)
int lp_init(fsd_ctx *ctx);
void lp_free(fsd_ctx *ctx);
int64_t lp_nparam(fsd_ctx *ctx);
int lp_set_weights(fsd_ctx *ctx, const float *h_w, int64_t n);
int lp_get_weights(fsd_ctx *ctx, float *h_w, int64_t n, bool grad);
int lp_adam_state(fsd_ctx *ctx, float *h_m, float *h_v, int64_t n, bool set);
int lp_forward(fsd_ctx *ctx, const float4 *pos, bool train, int inject, int *failed);
void lp_last_factor(const fsd_ctx *ctx, float *lambda, int *shrunk, int *plain);
int lp_apply(fsd_ctx *ctx, const float *b0, const float *b1, const float *bsaddle, float *x, int mode);
int lp_loss_grad(fsd_ctx *ctx, int nsamp, const float *d_x, const float *d_b, double *loss, bool grad);
int lp_adam(fsd_ctx *ctx, float lr, int step);
int lp_revert(fsd_ctx *ctx);
int lp_factor_loss_grad(fsd_ctx *ctx, double *loss, bool grad);
bool lp_active(const fsd_ctx *ctx);
bool lp_have_weights(const fsd_ctx *ctx);
void lp_set_active(fsd_ctx *ctx, bool on);
int lp_export(fsd_ctx *ctx, float *h_oute, float *h_outn, float *h_W, float *h_Dinv);
int lp_levels(fsd_ctx *ctx, int32_t *h_levels, int64_t n);


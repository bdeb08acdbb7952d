/*
 * fsd.h - C ABI of the MI355X Fast Stokesian Dynamics velocity/integration step.
 *
 * This is the boundary the host side (the C++ `Stokes` integrator and the
 * Python `integrate.PSEv3` class) calls into.  It replaces the reference's
 * host->device seam `Stokes_StepOne` (Stokes.cuh:81-108) and the setup calls
 * of `class Stokes` (Stokes.h:43-101); the operator-level entry points below
 * expose the individual functions of that step so they can be tested and
 * timed on their own.
 *
 * Conventions
 *   - every array argument named d_* is a device pointer (HIP device memory)
 *     owned by the caller; h_* arguments are host pointers;
 *   - particle arrays use the reference layouts:
 *       pos, pos_gb     float4[N]  (x, y, z, w)             (HOOMD positions)
 *       ori             float[3N]  unit director             (Stokes.cc:1234)
 *       image           int[3N]                              (HOOMD images)
 *       force/torque    float[6N]  [Fx Fy Fz Tx Ty Tz]      (AppliedForce)
 *       velocity        float[11N] [U,W (6N) | E or S (5N)] (Velocity)
 *       saddle vectors  float[17N] [F,T (6N) | S (5N) | U,W (6N)];
 *   - return value 0 = success; negative = error (fsd_last_error has text);
 *     FSD_W_NOT_CONVERGED (> 0) = a solve stopped at its iteration limit;
 *   - each call is enqueued on the context's stream and returns after that
 *     stream has drained, so device buffers may be reused on return.
 */
#ifndef FSD_H
#define FSD_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define FSD_OK 0
#define FSD_E_PARAM (-1)          /* invalid parameter (setParams' exit() cases) */
#define FSD_E_HIP (-2)            /* HIP / rocFFT runtime failure (gpuErrchk) */
#define FSD_E_CAPACITY (-3)       /* a neighbour or sparse capacity was exceeded */
#define FSD_E_ICHOL (-4)          /* IC(0) failed after the relaxer limit */
#define FSD_E_STATE (-5)          /* call order violated (e.g. no nlist yet) */
#define FSD_W_NOT_CONVERGED 1     /* GMRES hit its iteration limit (Solvers.cu:121-125) */
#define FSD_W_PIVOT 2             /* the learned factor has a non-positive pivot (no loss computed) */

typedef struct fsd_ctx fsd_ctx;

/* Constructor arguments of Stokes (Stokes.cc:47-83, integrate.py:40-55) plus
 * the box and the knobs the reference hard-codes (Stokes.cc:1505-1506,
 * Solvers.cu:105-117, integrate.py:86). */
typedef struct fsd_params {
    int32_t N;                       /* particles in the group */
    float Lx, Ly, Lz, xy, xz, yz;    /* periodic box (HOOMD BoxDim) */
    float T;                         /* kT of the bath */
    uint32_t seed;
    float xi;                        /* Ewald splitting parameter */
    float error;                     /* PSE error tolerance */
    float max_strain;                /* box strain bound for the spreading support */
    float dt;
    int32_t period;                  /* stresslet output period (0 = never) */
    float ndsr, kappa, k_n, beta_AF, epsq;  /* interparticle force law */
    float sqm_B1, sqm_B2;            /* squirmer modes */
    uint32_t N_mix;
    float coef_B1_mask, coef_B2_mask, rot_diff, T_ext, omega_ext;
    /* not settable in the reference (defaults in brackets) */
    float gmres_tol;                 /* <= 0: = error [error] */
    int32_t gmres_restart;           /* <= 0: 50 */
    int32_t gmres_maxiter;           /* <= 0: 1000 */
    float rlub;                      /* <= 0: 4.0 */
    float rp;                        /* <= 0: 2.1 */
    float r_buff;                    /* < 0: 0.4 */
    int32_t parity;                  /* 1: reproduce the reference's k-vector constant (Helper_Mobility.cu:389-391) */
    float gmres_true_tol;            /* > 0: also require ||b - A x|| / ||b|| <= this (extra mat-vecs) */
    int32_t guess;                   /* initial guess of a step's solve: 0 = the previous step's solution
                                        (the reference, Integrator.cu:777); 1 = linear extrapolation
                                        2 x_{n-1} - x_{n-2} of the last two steps' solutions; 2 = quadratic
                                        extrapolation 3 x_{n-1} - 3 x_{n-2} + x_{n-3} (lower orders while
                                        fewer solutions exist) */
    int32_t gs_passes;               /* classical Gram-Schmidt passes per Arnoldi step: 2 (default, <= 0)
                                        or 1 (half the vector traffic, less orthogonal basis) */
    int32_t precond;                 /* factor of the saddle preconditioner G = P^T L^-T L^-1 P and how it is
                                        applied: 0 = IC(0), exact triangular solves (the reference,
                                        Precondition.cu:1802-1890); 1 = IC(0), jacobi_iters Jacobi iterations
                                        per triangular solve; 2 = learned factor (fsd_learned_set_weights),
                                        jacobi_iters block Jacobi iterations per triangular solve, IC(0)
                                        as its fallback (applied as fallback_jacobi says) */
    int32_t jacobi_iters;            /* Jacobi iterations per triangular solve in modes 1 and 2 (<= 0: 3) */
    int32_t factor_sweeps;           /* IC(0) by the level-scheduled recurrence (0) or by this many parallel
                                        fixed-point sweeps (Chow & Patel 2015) from the block-diagonal start */
    int32_t refine;                  /* 1: FP64-residual iterative refinement of the saddle solve (the solution
                                        accumulated in FP64, r = b - A x with the lubrication product and the
                                        subtraction in FP64, FP32 GMRES for each correction) to the true
                                        residual gmres_true_tol; 0: plain FP32 GMRES (Solvers.cu:55-126) */
    int32_t fallback_jacobi;         /* mode 2: the IC(0) of calibration and fallback steps applied by this
                                        many Jacobi iterations per triangular solve (no sweep plans are
                                        built); 0: exact solves */
    int32_t autotune;                /* 1: the first fsd_step runs fsd_autotune at its positions */
} fsd_params;

/* Derived PSE parameters (Stokes.cc:131-261). */
typedef struct fsd_pse_info {
    int32_t Nx, Ny, Nz, P;
    float gaussm, eta, rcut, gridh_x, gridh_y, gridh_z;
    int32_t ewald_n;
    float self_x, self_y;
    float prefac, expfac, quadW;
} fsd_pse_info;

typedef struct fsd_solve_stats {
    int32_t iterations;              /* GMRES iterations (Arnoldi steps) */
    int32_t restarts;
    float residual;                  /* final preconditioned residual estimate / ||b|| */
    float true_residual;             /* ||b - A x|| / ||b||, if requested (< 0 otherwise) */
    float ichol_relaxer;             /* lambda actually used by IC(0) */
    int32_t ichol_attempts;
    int64_t nnz;                     /* preconditioner nonzeros (reference CSR count) */
    int64_t nlist_total;             /* entries in the neighbour list */
    int64_t lub_pairs;               /* lubrication pair records (r < rlub) */
    int64_t lower_blocks;            /* 6x6 blocks of the IC(0) factor below the diagonal */
    int32_t levels;                  /* level-schedule depth of the triangular solves */
    int32_t lanczos_ff;              /* Krylov size of the far-field Brownian sampler (T > 0) */
    int32_t lanczos_nf;              /* Krylov size of the near-field Brownian sampler (T > 0) */
    int32_t sweep_steps;             /* steps of the one-launch triangular sweeps (0: one launch per level) */
    float initial_residual;          /* ||b - A x0|| / ||b|| (true residual) or its preconditioned form */
    int32_t nlist_rebuilt;           /* fsd_step: 1 if the neighbour list was rebuilt this step (some particle
                                        moved r_buff / 2 since the last build, or the box changed) */
    int32_t precond_used;            /* the solve's factor: 0 IC(0), 1 the learned factor (precond mode 2) */
    int32_t fallback;                /* mode 2: 0 none, 1 the learned factor had a non-positive pivot (IC(0)
                                        used), 2 a learned solve of the step (its own or, in a Brownian step,
                                        a random-finite-difference solve) reached IC(0)'s iteration count for
                                        that solve and was finished with IC(0), which the rest of the step
                                        then used, 3 the step skipped the learned factor (IC(0) used) after
                                        fsd_learned_set_skip iteration fallbacks in a row */
} fsd_solve_stats;

/* ---- lifecycle: Stokes::Stokes + setParams + setSparseMath + AllocateWorkSpaces
 *      (Stokes.cc:47-83, 118-696, 937-1063; Stokes_SparseMath.cc:15-163) ---- */
int fsd_create(const fsd_params *p, fsd_ctx **out);
void fsd_destroy(fsd_ctx *ctx);
const char *fsd_last_error(const fsd_ctx *ctx);
int fsd_get_pse_info(const fsd_ctx *ctx, fsd_pse_info *out);

/* Stokes::setResistanceTable (Stokes_ResistanceTable.cc:34-23057):
 * h_dist[1000], h_vals[22000] in the reference's order. */
int fsd_set_resistance_table(fsd_ctx *ctx, const float *h_dist, const float *h_vals);
/* Stokes::setFriction (Stokes.cc:1132-1191): "NONE" (the default) or "TYPE1"
 * (tangential friction added to the resistance table for gaps h <= h0, with
 * coefficients h_alpha6[6]; parity = 1 writes the table columns the
 * reference's code writes); other types return FSD_E_PARAM. */
int fsd_set_friction(fsd_ctx *ctx, const char *type, float h0, const float *h_alpha6);
/* the resistance values in use after setFriction, h_vals[22000] (Stokes::m_ResTable_vals) */
int fsd_get_resistance_table(const fsd_ctx *ctx, float *h_vals);
/* Stokes::setT / set_params (Stokes.h:64-67, integrate.py:139-149) */
int fsd_set_temperature(fsd_ctx *ctx, float T);
/* box tilt for a deforming box (HOOMD box update between steps) */
int fsd_set_box(fsd_ctx *ctx, float Lx, float Ly, float Lz, float xy, float xz, float yz);
/* real-space Ewald table (Stokes.cc:344-678): h_tab[8*(ewald_n+1)] */
int fsd_get_ewald_table(const fsd_ctx *ctx, float *h_tab, int32_t capacity);

/* ---- the step: Stokes::integrateStepOne -> Stokes_StepOne
 *      (Stokes.cc:1197-1683, Stokes.cu:119-273) ----
 * d_force6: NULL -> the reference's interparticle force law
 *           (Stokes_SetForce_manually_kernel, Helper_Stokes.cu:65-172);
 *           otherwise the applied force/torque to use instead.
 * d_noise_ang3: rotational noise (NULL -> none)
 * d_vel11: output Velocity array (disturbance U,W then stresslet on output steps).
 * Positions, orientations, unwrapped positions and images are advanced in place. */
int fsd_step(fsd_ctx *ctx, uint32_t timestep, float shear_rate,
             float *d_pos4, float *d_ori3, float *d_pos_gb4, int32_t *d_image3,
             const float *d_force6, const float *d_noise_ang3, float *d_vel11,
             fsd_solve_stats *stats);

/* ---- operator-level entry points ---- */
/* HOOMD NeighborList::compute (Stokes.cc:1216): full list, r < rcut + r_buff. */
int fsd_build_nlist(fsd_ctx *ctx, const float *d_pos4);
/* copy the current list to the host: h_nneigh[N]; h_nbrs gets each particle's
 * neighbours ascending, packed; returns the total in *total. */
int fsd_get_nlist(fsd_ctx *ctx, int32_t *h_nneigh, int32_t *h_nbrs, int64_t capacity, int64_t *total);

/* Lubrication_RFU_kernel (Lubrication.cu:69-309): d_F6 = R_FU d_U6 */
int fsd_rfu(fsd_ctx *ctx, const float *d_pos4, const float *d_U6, float *d_F6);
/* Allocation-free variant of fsd_rfu.  It has no counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The lubrication pair
 * records go into a fixed device array of the library (one per GPU, held until
 * the process ends) instead of context buffers that fsd_rfu regrows.  Limits:
 * N <= FSD_NOALLOC_MAX_N (else FSD_E_PARAM) and at most FSD_NOALLOC_MAX_PAIRS
 * lubrication pairs per particle (else FSD_E_CAPACITY).  Needs a neighbour list
 * (fsd_build_nlist).  Calls that overlap on one GPU share the array and race. */
#define FSD_NOALLOC_MAX_N 8192
#define FSD_NOALLOC_MAX_PAIRS 48
int fsd_rfu_noalloc(fsd_ctx *ctx, const float *d_pos4, const float *d_U6, float *d_F6);
/* Lubrication_RFE_kernel (Lubrication.cu:350-608): d_F6 += R_FE : E */
int fsd_rfe(fsd_ctx *ctx, const float *d_pos4, float shear_rate, const float *d_ori3, float *d_F6);
/* Lubrication_RSU_kernel (Lubrication.cu:650-911): d_S5 -= R_SU d_U6 */
int fsd_rsu(fsd_ctx *ctx, const float *d_pos4, const float *d_U6, float *d_S5);
/* Lubrication_RSE_kernel (Lubrication.cu:953-1211): d_S5 += R_SE : E */
int fsd_rse(fsd_ctx *ctx, const float *d_pos4, float shear_rate, const float *d_ori3, float *d_S5);

/* Mobility_GeneralizedMobility (Mobility.cu:1488-1542): d_gU11 = M d_gF11;
 * parts: 1 real space only, 2 wave space only, 3 both. */
int fsd_mobility(fsd_ctx *ctx, const float *d_pos4, const float *d_gF11, float *d_gU11, int32_t parts);

/* Precondition_Wrap (Precondition.cu:1326-1497): prune, build R_FU^nf, RCM,
 * IC(0) with relaxer doubling.  Uses the current neighbour list. */
int fsd_build_precond(fsd_ctx *ctx, const float *d_pos4, fsd_solve_stats *stats);
/* ---- autotuner (no counterpart in the reference, whose launch shapes are fixed) ----
 * Times every variant of the kernels that compute the same operator at these positions
 * (spreading, contraction and x-transform kernels of the wave-space part; lanes per
 * particle and workgroup size of the real-space sum and of R_FU) and keeps the fastest.
 * h_choice8 (may be null): spread (0 tiled, 1 per-tile), contract (0 tiled, 1 per
 * particle), xpass (0 fused x transforms, 1 3-D rocFFT), real-space lanes, real-space
 * workgroup, R_FU lanes, R_FU workgroup, tuned; h_us3 (may be null): the best time per
 * apply in microseconds of the wave part, the real part and R_FU.  Allocates scratch. */
int fsd_autotune(fsd_ctx *ctx, const float *d_pos4, int32_t *h_choice8, float *h_us3);
int fsd_get_tuning(fsd_ctx *ctx, int32_t *h_choice8, float *h_us3);
/* sets the seven choices of fsd_autotune's h_choice8 (tests, or a stored tuning) */
int fsd_set_tuning(fsd_ctx *ctx, const int32_t *h_choice7);
/* Export of the preconditioner in the reference's formats:
 *   h_pn[N] pruned counts (self included), h_plist packed ascending (total = sum h_pn),
 *   h_perm[N] RCM permutation (new -> old),
 *   h_rowptr[6N+1], h_col[nnz], h_val[nnz]: R_FU^nf CSR before reordering (coo2csr of
 *     Precondition_RFU_kernel's COO), h_prow/h_pcol: pattern after csrperm,
 *   h_L[nnz]: IC(0) factor on the permuted pattern (upper part zero).
 * Any pointer may be NULL; capacities are in elements. */
int fsd_get_precond(fsd_ctx *ctx, int32_t *h_pn, int32_t *h_plist, int64_t plist_cap, int32_t *h_perm,
                    int32_t *h_rowptr, int32_t *h_col, float *h_val, int32_t *h_prow, int32_t *h_pcol,
                    float *h_L, int64_t nnz_cap, int64_t *nnz_out);
/* Precondition_Saddle_RFUmultiply (Precondition.cu:1802-1890): d_y6 = P^T L^-T L^-1 P d_x6 */
int fsd_precond_apply(fsd_ctx *ctx, const float *d_x6, float *d_y6);

/* Saddle_Multiply (Saddle.cu:80-151) */
int fsd_saddle_multiply(fsd_ctx *ctx, const float *d_pos4, const float *d_x17, float *d_y17);
/* Saddle_Preconditioner (Saddle.cu:171-264) */
int fsd_saddle_precond(fsd_ctx *ctx, const float *d_b17, float *d_x17);
/* Solvers_Saddle (Solvers.cu:55-126): CUSP-semantics left-preconditioned GMRES;
 * d_sol17 holds the initial guess on entry. want_true_residual adds one mat-vec. */
int fsd_saddle_solve(fsd_ctx *ctx, const float *d_pos4, const float *d_rhs17, float *d_sol17,
                     int32_t want_true_residual, fsd_solve_stats *stats);

/* Stokes_SetForce_manually_kernel (Helper_Stokes.cu:65-172) */
int fsd_force_law(fsd_ctx *ctx, const float *d_pos4, float T_ext, float *d_F6);

/* ---- Brownian motion (T > 0) ----
 * Noise: by default the library draws unit normals itself (counter-based
 * Philox-4x32-10 keyed by seed, timestep and stream), where the reference
 * draws cuRAND Philox streams seeded per step (Stokes.cc:1437-1441).
 * fsd_set_noise injects the unit normals instead (device pointers kept until
 * the next call; NULL = draw that stream again), which is how results are
 * compared with other implementations:
 *   d_ff_real11  11N  far-field real-space psi, generalised layout [F,T | S]
 *   d_ff_grid6M  6M   far-field grid noise, read at [q, q+6) for grid point q
 *                     (q = idx with parity = 1, the reference's reads; 6 idx otherwise)
 *   d_nf6        6N   near-field psi (scaled by sqrt(2T/dt) inside)
 *   d_rfd6       6N   random-finite-difference psi */
int fsd_set_noise(fsd_ctx *ctx, const float *d_ff_real11, const float *d_ff_grid6M, const float *d_nf6,
                  const float *d_rfd6);
/* Brownian_FarField_SlipVelocity (Brownian_FarField.cu:829-1061): d_out11 = U_B^ff
 * (wave space from the grid noise + sqrt(2T/dt) M_real^(1/2) psi by Lanczos);
 * needs the neighbour list. */
int fsd_brownian_farfield(fsd_ctx *ctx, const float *d_pos4, uint32_t timestep, float *d_out11);
/* Brownian_NearField_Force (Brownian_NearField.cu:407-464): d_out6 = R_FU^(1/2) psi,
 * psi ~ N(0, 2T/dt); needs the neighbour list. */
int fsd_brownian_nearfield(fsd_ctx *ctx, const float *d_pos4, uint32_t timestep, float *d_out6);
/* Integrator_RFD (Integrator.cu:549-718): d_out6 = (T/eps)(U(x + eps psi/2) - U(x - eps psi/2)),
 * eps = error; needs the neighbour list and the preconditioner. */
int fsd_rfd_drift(fsd_ctx *ctx, const float *d_pos4, uint32_t timestep, float *d_out6);
/* Integrator_ExplicitEuler_Shear_kernel (Integrator.cu:126-228) */
int fsd_euler_shear(fsd_ctx *ctx, float *d_pos4, float *d_ori3, float *d_pos_gb4, int32_t *d_image3,
                    const float *d_vel6, const float *d_mask1, const float *d_noise3,
                    float dt, float shear_rate);

/* Second-order Runge-Kutta stages, present in the reference but never called by
 * its step.  Stage 1 = Integrator_ExplicitEuler1_Shear_kernel (Integrator.cu:232-326):
 * pos_out = pos_in + u dt (no wrap, pos.w kept), director rotated over dt,
 * pos_gb advanced by u dt / 2.  Stage 2 = Integrator_RK_Shear_kernel
 * (Integrator.cu:329-438): position and director = coef1 * state1 + coef2 *
 * state2 (director renormalised), moved over coef3 * dt and wrapped (image
 * updated, pos.w written 0), pos_gb advanced by u dt / 2.  d_mask1 nullable (the
 * context's B1 mask), d_noise3 nullable. */
int fsd_rk2_stage1(fsd_ctx *ctx, const float *d_pos_in4, const float *d_ori_in3, float *d_pos_out4, float *d_ori_out3,
                   float *d_pos_gb4, const float *d_vel6, const float *d_mask1, const float *d_noise3, float dt,
                   float shear_rate);
int fsd_rk2_stage2(fsd_ctx *ctx, float coef1, const float *d_pos_in1_4, const float *d_ori_in1_3, float coef2,
                   const float *d_pos_in2_4, const float *d_ori_in2_3, float *d_pos_out4, float *d_ori_out3,
                   float *d_pos_gb4, const float *d_vel6, const float *d_mask1, const float *d_noise3,
                   int32_t *d_image3, float coef3, float dt, float shear_rate);

/* ---- timing of the hot kernels (HIP events on the context's stream) ---- */
/* enable: bit mask of kernel classes to time (bit k = k-th name below, -1 = all, 0 = off);
 * an event pair is recorded around every launch of a selected class */
int fsd_profile_enable(fsd_ctx *ctx, int32_t enable);
/* time one launch in `every` of each selected class (default 1: every launch), so that the
 * event records perturb a long run less; fsd_profile_get then reports the timed launches */
int fsd_profile_every(fsd_ctx *ctx, int32_t every);
/* name: "spread", "green", "contract", "fft", "realspace", "rfu", "trsv", "ichol",
 * "gmres_vec", "nlist", "precond_build", "sweep", the host wall time of the preconditioner's
 * host plan ("plan_host") and of its upload/build phase ("finish_host"), the forward block
 * Jacobi iterations of the learned factor ("ljit") or of IC(0) ("jacobi"; one event pair
 * around the k iterations of an application, counted as k launches), and the learned
 * factor's network and assembly ("learned_net"); returns launches (phases) and total
 * milliseconds */
int fsd_profile_get(fsd_ctx *ctx, const char *name, int64_t *launches, double *total_ms);
int fsd_profile_reset(fsd_ctx *ctx);

/* ---- learned preconditioner factor (precond mode 2) ----
 * No counterpart in the reference: these replace the IC(0) factorisation of the
 * saddle preconditioner (Precondition.cu:1116-1307, 1448-1489) with a factor that a
 * message-passing network predicts on IC(0)'s own pattern (Precondition.cu:530-836).
 * The factor is block LDL^T, (D + W) D^-1 (D + W)^T, W on the pattern's lower pair
 * blocks and D one 6x6 pivot per particle (positive definite by construction); a step
 * falls back to IC(0) when the network's outputs are not finite, and a solve finishes
 * with IC(0) when it reaches the iteration count IC(0) needed for the same solve at the
 * last calibration step (every fsd_learned_set_calibration steps, default 50, one step
 * runs IC(0)).
 * Parameters are FP32 (fsd_learned_nparams of them); the network runs them in BF16. */
int fsd_learned_nparams(fsd_ctx *ctx, int64_t *n);
int fsd_learned_set_weights(fsd_ctx *ctx, const float *h_w, int64_t n);
int fsd_learned_get_weights(fsd_ctx *ctx, float *h_w, int64_t n);
/* gradient of the last fsd_learned_loss_grad with want_grad != 0 */
int fsd_learned_get_grad(fsd_ctx *ctx, float *h_g, int64_t n);
/* Adam moments (first, second), nparams each, so training can move between contexts */
int fsd_learned_get_adam(fsd_ctx *ctx, float *h_m, float *h_v, int64_t n);
int fsd_learned_set_adam(fsd_ctx *ctx, const float *h_m, const float *h_v, int64_t n);
int fsd_learned_set_calibration(fsd_ctx *ctx, int32_t every);
/* on (default): in fsd_step, a step that does not calibrate keeps the last computed
 * ordering of the particles (the reference's RCM of an earlier step's pruned graph,
 * Precondition.cu:1009) instead of computing the RCM of its own graph, which any ordering
 * factors validly, as long as the neighbour list that ordering was computed under has not
 * been rebuilt (no particle moved r_buff / 2 since).  Off: every step orders as the
 * reference does.  No counterpart in the reference. */
int fsd_learned_set_keep_order(fsd_ctx *ctx, int32_t on);
/* after this many iteration fallbacks in a row (fallback 2) the steps up to the next
 * calibration step use IC(0) without building the learned factor (fallback 3); the
 * calibration step resets the count (default 2; 0: never skip).  No counterpart in the
 * reference, which has no learned factor. */
int fsd_learned_set_skip(fsd_ctx *ctx, int32_t after);
/* the next learned factor gets a non-finite pivot shift at this particle, which makes
 * the build fall back to IC(0) (failure-path tests) */
int fsd_learned_inject_pivot(fsd_ctx *ctx, int32_t particle);
/* h_out6: steps, pivot fallbacks, iteration fallbacks, calibration steps, and of the
 * last learned factor the pivots built with a reduced fill term (safeguard level 1-3)
 * and those replaced by A_PP + lambda S (level 4) */
int fsd_learned_stats(fsd_ctx *ctx, int64_t *h_out6);
/* network outputs (32 per lower block, 32 per particle), pair blocks W (36 per lower
 * block) and inverse pivots D^-1 (36 per particle) of the active learned factor, in
 * the pattern's order (fsd_get_precond's permuted lower blocks); any may be null */
int fsd_learned_export(fsd_ctx *ctx, float *h_oute, float *h_outn, float *h_W, float *h_Dinv, int64_t edge_cap);
/* pivot safeguard level (0-3: fill scaled by 2^-level; 4: the uncorrected pivot) of every
 * particle in each of the 4 passes of the last learned factor, [pass][particle] in the
 * pattern's order; n = 4 N */
int fsd_learned_levels(fsd_ctx *ctx, int32_t *h_levels, int64_t n);
/* Training loss for nsamp samples: d_x17 holds x_s, d_b17 holds b_s = A x_s (A the saddle
 * operator, fsd_saddle_multiply), 17N each, consecutive.  With g_s = G (b_F - b_U)_s the
 * learned preconditioner's G (jacobi_iters iterations each way), the loss is
 * sum_s |g_s - (b_F - x_F)_s|^2 + |g_s - x_U,s|^2 over sum_s |x_s|^2: the part of
 * |P(A x) - x|^2 / |x|^2 that depends on the factor, P the saddle preconditioner.
 * Builds the pattern at d_pos4 (needs the neighbour list); want_grad != 0 also forms
 * the gradient (fsd_learned_get_grad).  Returns FSD_W_PIVOT on a non-positive pivot. */
int fsd_learned_loss_grad(fsd_ctx *ctx, const float *d_pos4, int32_t nsamp, const float *d_x17, const float *d_b17,
                          int32_t want_grad, double *loss);
/* Supervised factor loss: the learned factor's pivots D_P and pair blocks W_PQ against
 * IC(0)'s own block LDL^T on the same pattern (L_PP L_PP^T and L_PQ L_QQ^T), as the
 * relative squared errors averaged over particles plus those averaged over pairs
 * (pair errors relative to the pair block A).  Builds the pattern and IC(0) at d_pos4;
 * want_grad != 0 also forms the gradient.  Returns FSD_W_PIVOT on a non-positive pivot. */
int fsd_learned_factor_loss(fsd_ctx *ctx, const float *d_pos4, int32_t want_grad, double *loss);
/* one Adam step (Kingma & Ba 2015; beta 0.9, 0.999) on the last gradient */
int fsd_learned_adam(fsd_ctx *ctx, float lr, int32_t step);
/* undo the last Adam step (parameters and moments), e.g. after it produced a non-positive pivot */
int fsd_learned_revert(fsd_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* FSD_H */

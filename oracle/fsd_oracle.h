/*
 * FP64 CPU oracle of the Fast Stokesian Dynamics velocity step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates the reference algorithm
 * (GeZhouyang/FSD "PSEv3", the .cu and .cc files under source/) in plain C so that the HIP
 * library can be checked against it.  Only the tests, the smoke check and the
 * benchmark's CPU-baseline leg may load it; the product never does.
 *
 * Conventions
 *   - arithmetic in double; the few decisions the reference takes in float
 *     (minimum-image distances, cutoff tests, table indices, Gaussian stencil
 *     origins) are taken in float exactly as the reference, so that integer
 *     results (neighbour sets, sparsity patterns, permutations) are bit-equal;
 *   - particle arrays: pos float4[N] (x,y,z,w), generalised vectors flat
 *     double arrays in the reference layouts
 *       force/velocity 6N:   [Fx Fy Fz Tx Ty Tz] per particle
 *       stresslet 5N:        [Sxx Sxy Sxz Syz Syy] per particle
 *       saddle vector 17N:   [ F,T (6N) | S (5N) | U,W (6N) ];
 *   - every function returns 0 on success, a negative code on error.
 */
#ifndef FSD_ORACLE_H
#define FSD_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_box { float Lx, Ly, Lz, xy, xz, yz; } orc_box;

typedef struct orc_pse {
    int Nx, Ny, Nz, P;
    float gaussm, eta, xi, error, rcut;
    float gridh[3];
    int ewald_n;          /* table holds ewald_n+1 nodes */
    float ewald_dr;
    float self[2];
    float prefac, expfac, quadW;
} orc_pse;

typedef struct orc_lub {
    const float *dist;    /* 1000 */
    const float *vals;    /* 22000 */
    float table_min, table_dr;
    float rlub, rp;
} orc_lub;

/* ---- box (HOOMD BoxDim semantics, float) ---- */
void orc_min_image(const orc_box *b, float d[3]);
void orc_make_fraction(const orc_box *b, const float p[3], float f[3]);
void orc_wrap(const orc_box *b, float p[3], int img[3]);

/* ---- parameters (Stokes.cc:118-261) ---- */
int orc_pse_params(const orc_box *b, float xi, float error, float max_strain, orc_pse *out);
/* Direct Ewald sum of the FTS grand mobility (no PSE): real space over all images
 * within r_cut with the functions in FP64, wave space over all |k| <= k_cut */
void orc_mobility_direct(int N, const float *pos4, const orc_box *b, double xi, double r_cut, double k_cut,
                         const double *gF, double *gU, int parts);
/* OpenMP threads the oracle runs on */
int orc_threads(void);
/* gridk: M float4 (kx, ky, kz, w), parity!=0 reproduces the 3.1416926536 factor */
void orc_gridk(const orc_box *b, const orc_pse *p, int parity, double *gridk4);

/* ---- neighbour list: all j != i with |r_ij| < rlist (float min-image), sorted ---- */
int orc_nlist(int N, const float *pos4, const orc_box *b, float rlist,
              int *nneigh, int *head, int *nbrs, int cap);

/* ---- lubrication (Lubrication.cu) ---- */
void orc_rfu(int N, const float *pos4, const orc_box *b, const int *nneigh, const int *head,
             const int *nbrs, const orc_lub *t, const double *U6, double *F6);
void orc_rfe(int N, const float *pos4, const orc_box *b, const int *nneigh, const int *head,
             const int *nbrs, const orc_lub *t, double shear_rate, double B2, const float *mask2,
             const float *ori3, double *F6_accum);
void orc_rsu(int N, const float *pos4, const orc_box *b, const int *nneigh, const int *head,
             const int *nbrs, const orc_lub *t, const double *U6, double *S5_accum);
void orc_rse(int N, const float *pos4, const orc_box *b, const int *nneigh, const int *head,
             const int *nbrs, const orc_lub *t, double shear_rate, double B2, const float *mask2,
             const float *ori3, double *S5_accum);

/* ---- far-field mobility (Mobility.cu) ---- */
/* generalised mobility: gF 11N (F,T | S) -> gU 11N (U,W | E5) */
/* far-field Brownian slip, wave-space part: noise 6M floats -> gU 11N
 * (Brownian_FarField.cu:113-395, 940-1011; serial, deterministic Nyquist writes) */
void orc_bro_ff_wave(int N, const float *pos4, const orc_box *b, const orc_pse *p, const double *gk,
                     const float *noise, int parity, double T, double dt, double *gU);
void orc_mobility(int N, const float *pos4, const orc_box *b, const orc_pse *p, const float *ewald_tab,
                  const double *gridk4, const int *nneigh, const int *head, const int *nbrs,
                  const double *gF, double *gU, int parts /* 1 real, 2 wave, 3 both */);

/* ---- preconditioner (Precondition.cu) ---- */
typedef struct orc_csr {
    int n, nnz;
    int *rowptr, *col;
    double *val;
} orc_csr;
/* pruned list (self included, ascending) ; returns total entries */
int orc_prune(int N, const float *pos4, const orc_box *b, const int *nneigh, const int *head,
              const int *nbrs, float rp, int *pn, int *phead /*N+1*/, int *plist, int *pless, int cap);
/* R_FU^nf on the pruned pattern in the reference's CSR order (row-sorted COO) */
int orc_build_rfu_csr(int N, const float *pos4, const orc_box *b, const int *pn, const int *phead,
                      const int *plist, const int *pless, const orc_lub *t, orc_csr *A);
/* SPARSPAK GENRCM on the pruned adjacency (self loops included); perm[new] = old */
void orc_rcm(int N, const int *phead, const int *plist, int *perm);
/* B = A(p,p) with p[6i+c] = 6 perm[i] + c, columns sorted within rows */
int orc_permute(const orc_csr *A, const int *perm, int N, orc_csr *B);
/* IC(0) in place on B (lower triangle used), with diag += lambda*(1 | 4/3);
 * returns 0, or 1 on a non-positive pivot */
int orc_ichol(orc_csr *B, double lambda);
/* G v = P^T L^-T L^-1 P v */
void orc_precond_apply(const orc_csr *L, const int *perm, int N, const double *x, double *y);
void orc_csr_free(orc_csr *A);

/* ---- saddle point (Saddle.cu) ---- */
typedef struct orc_system {
    int N;
    const float *pos4;
    orc_box box;
    const orc_pse *pse;
    const float *ewald_tab;
    const double *gridk4;
    const int *nneigh, *head, *nbrs;
    const orc_lub *lub;
    const orc_csr *L;      /* may be NULL (no preconditioner) */
    const int *perm;
} orc_system;
void orc_saddle_multiply(const orc_system *s, const double *x, double *y);
void orc_saddle_precond(const orc_system *s, const double *b, double *x);
/* CUSP-semantics left-preconditioned GMRES(restart); x in/out; returns iterations */
int orc_gmres(const orc_system *s, const double *b, double *x, int restart, int maxit, double tol,
              int use_precond, double *resid_out);

/* ---- forces and integration ---- */
void orc_force_law(int N, const float *pos4, const orc_box *b, const int *nneigh, const int *head,
                   const int *nbrs, double ndsr, double k_n, double kappa, double beta, double epsq,
                   double T_ext, double *F6);
void orc_euler_shear(int N, float *pos4, float *ori3, float *pos_gb4, int *image3, const orc_box *b,
                     const double *vel6, double B1, const float *mask1, const float *noise3,
                     double dt, double shear_rate);
void orc_rk2_stage1(int N, const float *pos_in4, const float *ori_in3, float *pos_out4, float *ori_out3,
                    float *pos_gb4, const double *vel6, double B1, const float *mask1, const float *noise3,
                    double dt, double shear_rate);
void orc_rk2_stage2(int N, double c1, const float *pos1, const float *ori1, double c2, const float *pos2,
                    const float *ori2, float *pos_out4, float *ori_out3, float *pos_gb4, const double *vel6,
                    double B1, const float *mask1, const float *noise3, int *image3, const orc_box *b, double c3,
                    double dt, double shear_rate);

#ifdef __cplusplus
}
#endif
#endif

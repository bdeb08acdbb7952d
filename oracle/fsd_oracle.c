/*
 * FP64 CPU oracle of the FSD velocity step.  TEST INFRASTRUCTURE ONLY: loaded
 * by tests/, the smoke check and the benchmark's CPU-baseline leg, never by
 * the product.  Each function cites the reference lines it restates
 * (paths relative to the reference's source/ directory).
 *
 * Float-faithful pieces (cutoff tests, table indices, stencil origins) are
 * compiled without floating-point contraction so that they round exactly
 * like the library's kernels; everything else is double.
 */
#include "fsd_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#pragma STDC FP_CONTRACT OFF

void fsd_ewald_self(float xi, float self_out[2]);  /* shared table code */
void fsd_ewald_functions(double r, double xi, double out[7]);

/* ------------------------------------------------------------------ box */
/* HOOMD BoxDim::minImage (device form, rint rounding). */
void orc_min_image(const orc_box *b, float d[3])
{
    const float iLx = 1.0f / b->Lx, iLy = 1.0f / b->Ly, iLz = 1.0f / b->Lz;
    float img = rintf(d[2] * iLz);
    float t = b->Lz * b->xz; t = t * img; d[0] = d[0] - t;
    t = b->Lz * b->yz; t = t * img; d[1] = d[1] - t;
    t = b->Lz * img; d[2] = d[2] - t;
    img = rintf(d[1] * iLy);
    t = b->Ly * b->xy; t = t * img; d[0] = d[0] - t;
    t = b->Ly * img; d[1] = d[1] - t;
    img = rintf(d[0] * iLx);
    t = b->Lx * img; d[0] = d[0] - t;
}

/* HOOMD BoxDim::makeFraction. */
void orc_make_fraction(const orc_box *b, const float p[3], float f[3])
{
    float dx = p[0] + 0.5f * b->Lx, dy = p[1] + 0.5f * b->Ly, dz = p[2] + 0.5f * b->Lz;
    float t = b->yz * b->xy; t = b->xz - t; t = t * p[2];
    float u = b->xy * p[1]; t = t + u; dx = dx - t;
    u = b->yz * p[2]; dy = dy - u;
    f[0] = dx / b->Lx; f[1] = dy / b->Ly; f[2] = dz / b->Lz;
}

/* HOOMD BoxDim::wrap (single image shift per axis, x then y then z). */
void orc_wrap(const orc_box *b, float w[3], int img[3])
{
    const float hx = 0.5f * b->Lx, hy = 0.5f * b->Ly, hz = 0.5f * b->Lz;
    float tilt_x = (b->xz - b->xy * b->yz) * w[2] + b->xy * w[1];
    if (w[0] >= hx + tilt_x) { w[0] -= b->Lx; img[0]++; }
    else if (w[0] < -hx + tilt_x) { w[0] += b->Lx; img[0]--; }
    float tilt_y = b->yz * w[2];
    if (w[1] >= hy + tilt_y) { w[1] -= b->Ly; w[0] -= b->Ly * b->xy; img[1]++; }
    else if (w[1] < -hy + tilt_y) { w[1] += b->Ly; w[0] += b->Ly * b->xy; img[1]--; }
    if (w[2] >= hz) { w[2] -= b->Lz; w[1] -= b->Lz * b->yz; w[0] -= b->Lz * b->xz; img[2]++; }
    else if (w[2] < -hz) { w[2] += b->Lz; w[1] += b->Lz * b->yz; w[0] += b->Lz * b->xz; img[2]--; }
}

/* squared float distance of j relative to i, min-imaged; d gets the vector j - i */
static float pair_d2(const orc_box *b, const float *pi, const float *pj, float d[3])
{
    d[0] = pj[0] - pi[0]; d[1] = pj[1] - pi[1]; d[2] = pj[2] - pi[2];
    orc_min_image(b, d);
    float s = d[0] * d[0]; float t = d[1] * d[1]; s = s + t; t = d[2] * d[2]; s = s + t;
    return s;
}

/* ---------------------------------------------------------- parameters */
/* Stokes::setParams, Stokes.cc:131-261 and Mobility.cu:930-935. */
int orc_pse_params(const orc_box *b, float xi, float error, float max_strain, orc_pse *o)
{
    memset(o, 0, sizeof(*o));
    o->xi = xi; o->error = error;
    o->rcut = sqrtf(-logf(error)) / xi;
    int kmax = (int)(2.0 * sqrtf(-logf(error)) * xi) + 1;
    if (o->rcut > b->Lx / 2.0 || o->rcut > b->Ly / 2.0 || o->rcut > b->Lz / 2.0) return -1;
    int n[3];
    const float L[3] = {b->Lx, b->Ly, b->Lz};
    for (int d = 0; d < 3; ++d) n[d] = (int)(kmax * L[d] / (2.0 * 3.1415926536) * 2.0) + 1;
    /* round up to the next 2^a 3^b 5^c in [8, 512] */
    for (int d = 0; d < 3; ++d) {
        int best = 1 << 30;
        for (int p2 = 1; p2 <= 512; p2 *= 2)
            for (int p3 = 1; p2 * p3 <= 512; p3 *= 3)
                for (int p5 = 1; p2 * p3 * p5 <= 512; p5 *= 5) {
                    int m = p2 * p3 * p5;
                    if (m >= 8 && m >= n[d] && m < best) best = m;
                }
        if (best == (1 << 30)) return -2;
        n[d] = best;
    }
    o->Nx = n[0]; o->Ny = n[1]; o->Nz = n[2];
    if ((long)o->Nx * o->Ny * o->Nz > 512L * 512 * 512) return -2;
    float gamma = max_strain, gamma2 = gamma * gamma;
    float lambda = (float)(1.0 + gamma2 / 2.0 + gamma * sqrtf((float)(1.0 + gamma2 / 4.0)));
    o->gridh[0] = b->Lx / (float)o->Nx; o->gridh[1] = b->Ly / (float)o->Ny; o->gridh[2] = b->Lz / (float)o->Nz;
    float gm = 1.0f;
    while (erfcf(gm / sqrtf((float)(2.0 * lambda))) > error) gm = (float)(gm + 0.01);
    o->gaussm = gm;
    o->P = (int)(gm * gm / 3.1415926536) + 1;
    float w = (float)(o->P * o->gridh[0] / 2.0);
    float xisq = xi * xi;
    o->eta = (float)((2.0 * w / gm) * (2.0 * w / gm) * xisq);
    if (o->P > o->Nx || o->P > o->Ny || o->P > o->Nz) return -3;
    o->ewald_dr = 0.001f;
    o->ewald_n = (int)(o->rcut / o->ewald_dr - 1.0f);
    fsd_ewald_self(xi, o->self);
    o->quadW = o->gridh[0] * o->gridh[1] * o->gridh[2];
    o->prefac = (float)((2.0 * xisq / 3.1415926536 / o->eta) * sqrtf((float)(2.0 * xisq / 3.1415926536 / o->eta)));
    o->expfac = (float)(2.0 * xisq / o->eta);
    return 0;
}

/* Mobility_SetGridk_kernel, Helper_Mobility.cu:351-410 (double). */
void orc_gridk(const orc_box *b, const orc_pse *p, int parity, double *gk)
{
    const double twopi = 2.0 * (parity ? 3.1416926536 : 3.14159265358979323846);
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    const double xisq = (double)p->xi * p->xi;
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                long id = ((long)i * Ny + j) * Nz + k;
                double kx = (i < (Nx + 1) / 2) ? i : i - Nx;
                double ky = (((j < (Ny + 1) / 2) ? j : j - Ny) - (double)b->xy * kx * b->Ly / b->Lx) / b->Ly;
                double kz = ((k < (Nz + 1) / 2) ? k : k - Nz) / (double)b->Lz;
                kx = kx / b->Lx;
                kx *= twopi; ky *= twopi; kz *= twopi;
                double k2 = kx * kx + ky * ky + kz * kz;
                double w = 0.0;
                if (id != 0)
                    w = 6.0 * 3.1415926536 * (1.0 + k2 / 4.0 / xisq) * exp(-(1.0 - p->eta) * k2 / 4.0 / xisq)
                        / k2 / (double)((long)Nx * Ny * Nz);
                gk[4 * id] = kx; gk[4 * id + 1] = ky; gk[4 * id + 2] = kz; gk[4 * id + 3] = w;
            }
}

/* ------------------------------------------------------ neighbour list */
int orc_nlist(int N, const float *pos4, const orc_box *b, float rlist, int *nneigh, int *head,
              int *nbrs, int cap)
{
    const float rl2 = rlist * rlist;
    int tot = 0;
    for (int i = 0; i < N; ++i) {
        head[i] = tot;
        int c = 0;
        for (int j = 0; j < N; ++j) {
            if (j == i) continue;
            float d[3];
            if (pair_d2(b, pos4 + 4 * i, pos4 + 4 * j, d) < rl2) {
                if (tot < cap) nbrs[tot] = j;
                ++tot; ++c;
            }
        }
        nneigh[i] = c;
    }
    return tot;
}

/* ------------------------------------------------------------ lubrication */
/* table lookup of `ncol` functions starting at column c0, as the kernels
 * in Lubrication.cu:172-231 (contact row 232 for r <= 2.001). */
static void lub_lookup(const orc_lub *t, float dist, int c0, int ncol, int contact_row, double *out)
{
    if ((double)dist <= 2.001 && contact_row >= 0) {
        for (int c = 0; c < ncol; ++c) out[c] = t->vals[22 * contact_row + c0 + c];
        return;
    }
    float arg = (float)(((double)dist - 2.0) / (double)t->table_min);
    int ind = (int)(log10f(arg) / t->table_dr);
    if (ind >= 999) { /* band r in [~3.99968, 4): the last row is all zeros */
        for (int c = 0; c < ncol; ++c) out[c] = t->vals[22 * 999 + c0 + c];
        return;
    }
    double dl = t->dist[ind], du = t->dist[ind + 1];
    double fac = ((double)dist - dl) / (du - dl);
    for (int c = 0; c < ncol; ++c) {
        double lo = t->vals[22 * ind + c0 + c], hi = t->vals[22 * (ind + 1) + c0 + c];
        out[c] = lo + (hi - lo) * fac;
    }
}

static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
/* (r x v) with the reference's sign convention eps.r.v (Lubrication.cu:243-254) */
static inline void epsr(const double *r, const double *v, double *o)
{
    o[0] = r[2] * v[1] - r[1] * v[2];
    o[1] = -r[2] * v[0] + r[0] * v[2];
    o[2] = r[1] * v[0] - r[0] * v[1];
}

/* Lubrication_RFU_kernel, Lubrication.cu:69-309 */
void orc_rfu(int N, const float *pos4, const orc_box *b, const int *nneigh, const int *head,
             const int *nbrs, const orc_lub *t, const double *U, double *F)
{
    const float rl2 = t->rlub * t->rlub;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < N; ++i) {
        double fi[3] = {0, 0, 0}, li[3] = {0, 0, 0};
        const double *ui = U + 6 * i, *wi = U + 6 * i + 3;
        for (int n = 0; n < nneigh[i]; ++n) {
            int j = nbrs[head[i] + n];
            float d[3];
            float d2 = pair_d2(b, pos4 + 4 * i, pos4 + 4 * j, d);
            if (!(d2 < rl2 && d2 > 0.f)) continue;
            float dist = sqrtf(d2);
            double v[10];
            lub_lookup(t, dist, 0, 10, 232, v);
            double XA11 = v[0], XA12 = v[1], YA11 = v[2], YA12 = v[3], YB11 = v[4], YB12 = v[5];
            double XC11 = v[6], XC12 = v[7], YC11 = v[8], YC12 = v[9];
            double r[3] = {d[0] / (double)dist, d[1] / (double)dist, d[2] / (double)dist};
            const double *uj = U + 6 * j, *wj = U + 6 * j + 3;
            double rdui = dot3(r, ui), rduj = dot3(r, uj), rdwi = dot3(r, wi), rdwj = dot3(r, wj);
            double eui[3], ewi[3], euj[3], ewj[3];
            epsr(r, ui, eui); epsr(r, wi, ewi); epsr(r, uj, euj); epsr(r, wj, ewj);
            for (int a = 0; a < 3; ++a) {
                fi[a] += (XA11 - YA11) * rdui * r[a] + YA11 * ui[a] + (XA12 - YA12) * rduj * r[a] + YA12 * uj[a]
                         - YB11 * ewi[a] + YB12 * ewj[a];
                li[a] += YB11 * eui[a] + YB12 * euj[a] + (XC11 - YC11) * rdwi * r[a] + YC11 * wi[a]
                         + (XC12 - YC12) * rdwj * r[a] + YC12 * wj[a];
            }
        }
        for (int a = 0; a < 3; ++a) { F[6 * i + a] = fi[a]; F[6 * i + 3 + a] = li[a]; }
    }
}

/* strain-rate block of particle k: E_inf (2Exy = shear_rate) plus the B2
 * squirming term (Lubrication.cu:389-418); returns the symmetric 3x3 */
static void e_block(double shear_rate, double B2, const float *mask2, const float *ori3, int k, double E[3][3])
{
    double e[5] = {0, shear_rate, 0, 0, 0};
    if (B2 != 0.0 && mask2 && ori3) {
        double b2 = 0.6 * B2 * mask2[k];
        double px = ori3[3 * k], py = ori3[3 * k + 1], pz = ori3[3 * k + 2];
        e[0] += b2 * (px * px - pz * pz); e[1] += b2 * (2. * px * py); e[2] += b2 * (2. * px * pz);
        e[3] += b2 * (2. * py * pz); e[4] += b2 * (py * py - pz * pz);
    }
    E[0][0] = (2.0 * e[0] - e[4]) / 3.0; E[0][1] = 0.5 * e[1]; E[0][2] = 0.5 * e[2];
    E[1][0] = E[0][1]; E[1][1] = (-e[0] + 2.0 * e[4]) / 3.0; E[1][2] = 0.5 * e[3];
    E[2][0] = E[0][2]; E[2][1] = E[1][2]; E[2][2] = -(e[0] + e[4]) / 3.0;
}

static void matvec3(double E[3][3], const double *r, double *o)
{
    for (int a = 0; a < 3; ++a) o[a] = E[a][0] * r[0] + E[a][1] * r[1] + E[a][2] * r[2];
}

/* Lubrication_RFE_kernel, Lubrication.cu:350-608: F += R_FE : E */
void orc_rfe(int N, const float *pos4, const orc_box *b, const int *nneigh, const int *head,
             const int *nbrs, const orc_lub *t, double shear_rate, double B2, const float *mask2,
             const float *ori3, double *F)
{
    const float rl2 = t->rlub * t->rlub;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < N; ++i) {
        double Ei[3][3];
        e_block(shear_rate, B2, mask2, ori3, i, Ei);
        for (int n = 0; n < nneigh[i]; ++n) {
            int j = nbrs[head[i] + n];
            double Ej[3][3];
            e_block(shear_rate, B2, mask2, ori3, j, Ej);
            float d[3];
            float d2 = pair_d2(b, pos4 + 4 * i, pos4 + 4 * j, d);
            if (!(d2 < rl2)) continue;
            float dist = sqrtf(d2);
            double v[6];
            lub_lookup(t, dist, 10, 6, 232, v);
            double XG11 = v[0], XG12 = v[1], YG11 = v[2], YG12 = v[3], YH11 = v[4], YH12 = v[5];
            double r[3] = {d[0] / (double)dist, d[1] / (double)dist, d[2] / (double)dist};
            double Edri[3], Edrj[3], eEi[3], eEj[3];
            matvec3(Ei, r, Edri); matvec3(Ej, r, Edrj);
            double rEi = dot3(r, Edri), rEj = dot3(r, Edrj);
            epsr(r, Edri, eEi); epsr(r, Edrj, eEj);
            double XG21 = -XG12, YG21 = -YG12, YH21 = YH12;
            for (int a = 0; a < 3; ++a) {
                F[6 * i + a] += (XG11 - 2.0 * YG11) * (-rEi) * r[a] + 2.0 * YG11 * (-Edri[a])
                                + (XG21 - 2.0 * YG21) * (-rEj) * r[a] + 2.0 * YG21 * (-Edrj[a]);
                F[6 * i + 3 + a] += YH11 * (2.0 * eEi[a]) + YH21 * (2.0 * eEj[a]);
            }
        }
    }
}

/* Lubrication_RSU_kernel, Lubrication.cu:650-911: S -= R_SU U */
void orc_rsu(int N, const float *pos4, const orc_box *b, const int *nneigh, const int *head,
             const int *nbrs, const orc_lub *t, const double *U, double *S)
{
    const float rl2 = t->rlub * t->rlub;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < N; ++i) {
        double si[5] = {0, 0, 0, 0, 0};
        const double *ui = U + 6 * i, *wi = U + 6 * i + 3;
        for (int n = 0; n < nneigh[i]; ++n) {
            int j = nbrs[head[i] + n];
            float d[3];
            float d2 = pair_d2(b, pos4 + 4 * i, pos4 + 4 * j, d);
            if (!(d2 < rl2)) continue;
            float dist = sqrtf(d2);
            double v[6];
            lub_lookup(t, dist, 10, 6, 232, v);
            double XG11 = v[0], XG12 = v[1], YG11 = v[2], YG12 = v[3], YH11 = v[4], YH12 = v[5];
            double r[3] = {d[0] / (double)dist, d[1] / (double)dist, d[2] / (double)dist};
            const double *uj = U + 6 * j, *wj = U + 6 * j + 3;
            double rdui = dot3(r, ui), rduj = dot3(r, uj);
            /* component pairs (a,b) of the 5-vector: xx, xy, xz, yz, yy */
            static const int ca[5] = {0, 0, 0, 1, 1}, cb[5] = {0, 1, 2, 2, 1};
            double ewi[3], ewj[3];
            epsr(r, wi, ewi); epsr(r, wj, ewj);
            for (int c = 0; c < 5; ++c) {
                int a = ca[c], bb = cb[c];
                double third = (a == bb) ? 1.0 / 3.0 : 0.0;
                si[c] += XG11 * (r[a] * r[bb] - third) * rdui + XG12 * (r[a] * r[bb] - third) * rduj
                         + YG11 * (ui[a] * r[bb] + r[a] * ui[bb] - 2.0 * r[a] * r[bb] * rdui)
                         + YG12 * (uj[a] * r[bb] + r[a] * uj[bb] - 2.0 * r[a] * r[bb] * rduj);
                si[c] += YH11 * (r[a] * ewi[bb] + ewi[a] * r[bb]) + YH12 * (r[a] * ewj[bb] + ewj[a] * r[bb]);
            }
        }
        for (int c = 0; c < 5; ++c) S[5 * i + c] -= si[c];
    }
}

/* Lubrication_RSE_kernel, Lubrication.cu:953-1211: S += R_SE : E (cutoff r^2 < 16) */
void orc_rse(int N, const float *pos4, const orc_box *b, const int *nneigh, const int *head,
             const int *nbrs, const orc_lub *t, double shear_rate, double B2, const float *mask2,
             const float *ori3, double *S)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < N; ++i) {
        double Ei[3][3];
        e_block(shear_rate, B2, mask2, ori3, i, Ei);
        double Si[5] = {0, 0, 0, 0, 0};
        for (int n = 0; n < nneigh[i]; ++n) {
            int j = nbrs[head[i] + n];
            double Ej[3][3];
            e_block(shear_rate, B2, mask2, ori3, j, Ej);
            float d[3];
            float d2 = pair_d2(b, pos4 + 4 * i, pos4 + 4 * j, d);
            if (!((double)d2 < 16.0)) continue;
            float dist = sqrtf(d2);
            double v[6];
            lub_lookup(t, dist, 16, 6, 232, v);
            double XM11 = v[0], XM12 = v[1], YM11 = v[2], YM12 = v[3], ZM11 = v[4], ZM12 = v[5];
            double r[3] = {d[0] / (double)dist, d[1] / (double)dist, d[2] / (double)dist};
            double Edri[3], Edrj[3];
            matvec3(Ei, r, Edri); matvec3(Ej, r, Edrj);
            double rEi = dot3(r, Edri), rEj = dot3(r, Edrj);
            static const int ca[5] = {0, 0, 0, 1, 1}, cb[5] = {0, 1, 2, 2, 1};
            for (int c = 0; c < 5; ++c) {
                int a = ca[c], bb = cb[c];
                double third = (a == bb) ? 1.0 / 3.0 : 0.0;
                double delta = (a == bb) ? 1.0 : 0.0;
                Si[c] += 1.5 * XM11 * (r[a] * r[bb] - third) * rEi + 1.5 * XM12 * (r[a] * r[bb] - third) * rEj
                         + 0.5 * YM11 * (2.0 * r[a] * Edri[bb] + 2.0 * r[bb] * Edri[a] - 4.0 * rEi * r[a] * r[bb])
                         + 0.5 * YM12 * (2.0 * r[a] * Edrj[bb] + 2.0 * r[bb] * Edrj[a] - 4.0 * rEj * r[a] * r[bb])
                         + 0.5 * ZM11 * (2 * Ei[a][bb] + (delta + r[a] * r[bb]) * rEi - 2.0 * r[a] * Edri[bb] - 2.0 * r[bb] * Edri[a])
                         + 0.5 * ZM12 * (2 * Ej[a][bb] + (delta + r[a] * r[bb]) * rEj - 2.0 * r[a] * Edrj[bb] - 2.0 * r[bb] * Edrj[a]);
            }
        }
        for (int c = 0; c < 5; ++c) S[5 * i + c] += Si[c];
    }
}

/* --------------------------------------------------------- FFT (double) */
/* in-place complex FFT of length n (n = 2^a 3^b 5^c), stride s, sign -1
 * forward / +1 inverse, unnormalised; mixed-radix recursive decimation. */
static void fft1(double complex *x, int n, int s, int sign, double complex *tmp)
{
    if (n == 1) return;
    int r = (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : (n % 5 == 0) ? 5 : n;
    int m = n / r;
    /* split into r decimated sub-sequences stored contiguously in tmp */
    for (int q = 0; q < r; ++q)
        for (int k = 0; k < m; ++k) tmp[q * m + k] = x[(long)(k * r + q) * s];
    for (int q = 0; q < r; ++q) fft1(tmp + q * m, m, 1, sign, tmp + n);
    for (int k = 0; k < m; ++k)
        for (int p = 0; p < r; ++p) {
            int kk = k + p * m;
            double complex acc = 0;
            for (int q = 0; q < r; ++q) {
                double ang = sign * 2.0 * M_PI * (double)q * (double)kk / (double)n;
                acc += tmp[q * m + k] * (cos(ang) + I * sin(ang));
            }
            x[(long)kk * s] = acc;
        }
}

static void fft3(double complex *g, int Nx, int Ny, int Nz, int sign)
{
    int nmax = Nx > Ny ? (Nx > Nz ? Nx : Nz) : (Ny > Nz ? Ny : Nz);
#pragma omp parallel
    {
        double complex *line = malloc(sizeof(double complex) * nmax);
        double complex *tmp = malloc(sizeof(double complex) * nmax * 8);
#pragma omp for collapse(2)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j) {
                double complex *p = g + ((long)i * Ny + j) * Nz;
                for (int k = 0; k < Nz; ++k) line[k] = p[k];
                fft1(line, Nz, 1, sign, tmp);
                for (int k = 0; k < Nz; ++k) p[k] = line[k];
            }
#pragma omp for collapse(2)
        for (int i = 0; i < Nx; ++i)
            for (int k = 0; k < Nz; ++k) {
                for (int j = 0; j < Ny; ++j) line[j] = g[((long)i * Ny + j) * Nz + k];
                fft1(line, Ny, 1, sign, tmp);
                for (int j = 0; j < Ny; ++j) g[((long)i * Ny + j) * Nz + k] = line[j];
            }
#pragma omp for collapse(2)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                for (int i = 0; i < Nx; ++i) line[i] = g[((long)i * Ny + j) * Nz + k];
                fft1(line, Nx, 1, sign, tmp);
                for (int i = 0; i < Nx; ++i) g[((long)i * Ny + j) * Nz + k] = line[i];
            }
        free(line); free(tmp);
    }
}

/* ------------------------------------------------------------- mobility */
/* torque+stresslet -> couplet (Mobility_TS2C_kernel, Helper_Mobility.cu:294-336);
 * ts = [Tx Ty Tz | Sxx Sxy Sxz Syz Syy] */
static void ts2c(const double *T, const double *S, double C[8])
{
    double Lx = -T[0], Ly = -T[1], Lz = -T[2];
    double Sxx = S[0], Sxy = S[1], Sxz = S[2], Syz = S[3], Syy = S[4];
    C[0] = Sxx; C[1] = Sxy + 0.5 * Lz; C[2] = Sxz - 0.5 * Ly; C[3] = Syz + 0.5 * Lx;
    C[4] = Syy; C[5] = Sxy - 0.5 * Lz; C[6] = Sxz + 0.5 * Ly; C[7] = Syz - 0.5 * Lx;
}

/* velocity gradient -> angular velocity + 5-component strain rate
 * (Mobility_D2WE_kernel, Helper_Mobility.cu:235-283) */
static void d2we(const double D[8], double W[3], double E5[5])
{
    W[0] = 0.5 * (D[3] - D[7]); W[1] = 0.5 * (D[6] - D[2]); W[2] = 0.5 * (D[1] - D[5]);
    double e0 = D[0], e1 = 0.5 * (D[1] + D[5]), e2 = 0.5 * (D[2] + D[6]), e3 = 0.5 * (D[3] + D[7]), e4 = D[4];
    E5[0] = 2 * e0 + e4; E5[1] = 2 * e1; E5[2] = 2 * e2; E5[3] = 2 * e3; E5[4] = 2 * e4 + e0;
}

/* Real-space FTS sum for particle i (Mobility_RealSpace_kernel, Mobility.cu:1046-1226). */
static void real_space_i(int i, const float *pos4, const orc_box *b, const orc_pse *p, const float *tab,
                         const int *nneigh, const int *head, const int *nbrs, const double *F3,
                         const double *C8, double u[3], double D[8])
{
    const double sx = p->self[0], sy = p->self[1];
    const double *F = F3 + 3 * i, *C = C8 + 8 * i;
    u[0] = sx * F[0]; u[1] = sx * F[1]; u[2] = sx * F[2];
    /* D layout: {xx', ., ., ., yy', ., ., .} as the reference's two float4 */
    D[0] = sy * (C[0] - 4. * C[0]); D[1] = sy * (C[1] - 4. * C[5]); D[2] = sy * (C[2] - 4. * C[6]); D[3] = sy * (C[3] - 4. * C[7]);
    D[4] = sy * (C[4] - 4. * C[4]); D[5] = sy * (C[5] - 4. * C[1]); D[6] = sy * (C[6] - 4. * C[2]); D[7] = sy * (C[7] - 4. * C[3]);
    const float mind2 = p->ewald_dr * p->ewald_dr, maxd2 = p->rcut * p->rcut;
    for (int n = 0; n < nneigh[i]; ++n) {
        int j = nbrs[head[i] + n];
        /* the reference uses r = x_i - x_j here */
        float d[3] = {pos4[4 * i] - pos4[4 * j], pos4[4 * i + 1] - pos4[4 * j + 1], pos4[4 * i + 2] - pos4[4 * j + 2]};
        orc_min_image(b, d);
        float d2 = d[0] * d[0]; float tt = d[1] * d[1]; d2 = d2 + tt; tt = d[2] * d[2]; d2 = d2 + tt;
        if (!(d2 < maxd2 && d2 >= mind2)) continue;
        float dist = sqrtf(d2);
        float xnum = (float)p->ewald_n * (dist - p->ewald_dr);
        int r_ind = (int)floorf(xnum / (p->rcut - p->ewald_dr));
        if (r_ind > p->ewald_n - 1) r_ind = p->ewald_n - 1;   /* defined clamp (reference reads past the table) */
        double fac = (double)dist / (double)p->ewald_dr - r_ind - 1.0;
        const float *m0 = tab + 8 * r_ind, *m1 = tab + 8 * (r_ind + 1);
        double f1 = m0[0] + (m1[0] - (double)m0[0]) * fac;
        double f2 = m0[1] + (m1[1] - (double)m0[1]) * fac;
        double g1 = -(m0[2] + (m1[2] - (double)m0[2]) * fac);
        double g2 = -(m0[3] + (m1[3] - (double)m0[3]) * fac);
        double h1 = m0[4] + (m1[4] - (double)m0[4]) * fac;
        double h2 = m0[5] + (m1[5] - (double)m0[5]) * fac;
        double h3 = m0[6] + (m1[6] - (double)m0[6]) * fac;
        double R[3] = {d[0] / (double)dist, d[1] / (double)dist, d[2] / (double)dist};
        const double *Fj = F3 + 3 * j, *Cj = C8 + 8 * j;
        double rdotf = dot3(R, Fj);
        /* full 3x3 couplet of j from the 8 stored components */
        double Cm[3][3] = {{Cj[0], Cj[1], Cj[2]}, {Cj[5], Cj[4], Cj[3]}, {Cj[6], Cj[7], -(Cj[0] + Cj[4])}};
        double Cdr[3], rdC[3];
        for (int a = 0; a < 3; ++a) {
            Cdr[a] = Cm[a][0] * R[0] + Cm[a][1] * R[1] + Cm[a][2] * R[2];
            rdC[a] = R[0] * Cm[0][a] + R[1] * Cm[1][a] + R[2] * Cm[2][a];
        }
        double rCr = dot3(R, Cdr);
        for (int a = 0; a < 3; ++a) {
            u[a] += f1 * Fj[a] + (f2 - f1) * rdotf * R[a];
            u[a] += g1 * (Cdr[a] - rCr * R[a]) + g2 * (rdC[a] - 4. * rCr * R[a]);
        }
        /* gradient entries in the stored order; (a,b) pairs and the
         * couplet component subtracted in the h1/h3 terms */
        static const int A[8] = {0, 0, 0, 1, 1, 1, 2, 2};
        static const int B[8] = {0, 1, 2, 2, 1, 0, 0, 1};
        static const int Cself[8] = {0, 1, 2, 3, 4, 5, 6, 7};
        static const int Cswap[8] = {0, 5, 6, 7, 4, 1, 2, 3};
        for (int c = 0; c < 8; ++c) {
            int a = A[c], bb = B[c];
            double del = (a == bb) ? 1.0 : 0.0;
            D[c] += -g1 * (R[a] * Fj[bb] - rdotf * R[a] * R[bb]) - g2 * (del * rdotf + Fj[a] * R[bb] - 4. * rdotf * R[a] * R[bb]);
            D[c] += h1 * (Cj[Cself[c]] - 4. * Cj[Cswap[c]]) + h2 * (R[a] * Cdr[bb] - rCr * R[a] * R[bb])
                    + h3 * (del * rCr + Cdr[a] * R[bb] + R[a] * rdC[bb] + rdC[a] * R[bb] - 6. * rCr * R[a] * R[bb] - Cj[Cswap[c]]);
        }
    }
}

/* PSE wave-space part (gpu_stokes_Mwave_wrap and kernels, Mobility.cu:131-1020). */
/* per particle stencil origin, computed in float as the reference (Mobility.cu:203-225) */
static int *stencil_origins(int N, const float *pos4, const orc_box *b, const orc_pse *p)
{
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz, P = p->P;
    int *org = malloc(sizeof(int) * 3 * N);
    for (int i = 0; i < N; ++i) {
        float f[3];
        orc_make_fraction(b, pos4 + 4 * i, f);
        float fx = f[0] * (float)Nx, fy = f[1] * (float)Ny, fz = f[2] * (float)Nz;
        int x = (int)fx, y = (int)fy, z = (int)fz;
        int Pd2 = P / 2;
        org[3 * i] = x - Pd2 + 1 - (P % 2) * ((fx - (float)x) < 0.5f);
        org[3 * i + 1] = y - Pd2 + 1 - (P % 2) * ((fy - (float)y) < 0.5f);
        org[3 * i + 2] = z - Pd2 + 1 - (P % 2) * ((fz - (float)z) < 0.5f);
    }
    return org;
}

static void contract(int N, const float *pos4, const orc_box *b, const orc_pse *p, const int *org,
                     const double complex *g, double *u3, double *D8);

static void wave_space(int N, const float *pos4, const orc_box *b, const orc_pse *p, const double *gk,
                       const double *F3, const double *C8, double *u3, double *D8)
{
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz, P = p->P;
    const long M = (long)Nx * Ny * Nz;
    const double xisq = (double)p->xi * p->xi;
    const double prefac = pow(2.0 * xisq / M_PI / p->eta, 1.5);
    const double expfac = 2.0 * xisq / p->eta;
    double complex *g = calloc((size_t)11 * M, sizeof(double complex));
    int *org = stencil_origins(N, pos4, b, p);
    const double Lx = b->Lx, Ly = b->Ly, Lz = b->Lz;
    /* spread (serial: grid accumulation) */
    for (int i = 0; i < N; ++i) {
        double C[8]; memcpy(C, C8 + 8 * i, sizeof(C));
        const double *F = F3 + 3 * i;
        for (int tx = 0; tx < P; ++tx)
            for (int ty = 0; ty < P; ++ty)
                for (int tz = 0; tz < P; ++tz) {
                    int xi_ = org[3 * i] + tx, yi_ = org[3 * i + 1] + ty, zi_ = org[3 * i + 2] + tz;
                    xi_ = (xi_ < 0) ? xi_ + Nx : (xi_ > Nx - 1 ? xi_ - Nx : xi_);
                    yi_ = (yi_ < 0) ? yi_ + Ny : (yi_ > Ny - 1 ? yi_ - Ny : yi_);
                    zi_ = (zi_ < 0) ? zi_ + Nz : (zi_ > Nz - 1 ? zi_ - Nz : zi_);
                    double gy = (double)p->gridh[1] * yi_ - Ly / 2;
                    double gx = (double)p->gridh[0] * xi_ - Lx / 2 + (double)b->xy * gy;
                    double gz = (double)p->gridh[2] * zi_ - Lz / 2;
                    float d[3] = {(float)(gx - pos4[4 * i]), (float)(gy - pos4[4 * i + 1]), (float)(gz - pos4[4 * i + 2])};
                    /* image choice in float; displacement re-evaluated in double */
                    float d0[3] = {d[0], d[1], d[2]};
                    orc_min_image(b, d);
                    double dx = (gx - pos4[4 * i]) + (double)(d[0] - d0[0]);
                    double dy = (gy - pos4[4 * i + 1]) + (double)(d[1] - d0[1]);
                    double dz = (gz - pos4[4 * i + 2]) + (double)(d[2] - d0[2]);
                    double w = prefac * exp(-expfac * (dx * dx + dy * dy + dz * dz));
                    long id = ((long)xi_ * Ny + yi_) * Nz + zi_;
                    g[0 * M + id] += w * F[0]; g[1 * M + id] += w * F[1]; g[2 * M + id] += w * F[2];
                    /* couplet grids XX XY XZ YZ YY YX ZX ZY */
                    for (int c = 0; c < 8; ++c) g[(3 + c) * M + id] += w * C[c];
                }
    }
    for (int c = 0; c < 11; ++c) fft3(g + c * M, Nx, Ny, Nz, -1);
    /* Green's function scaling (Mobility_WaveSpace_Green_kernel, Mobility.cu:315-479) */
#pragma omp parallel for
    for (long id = 0; id < M; ++id) {
        const double *tk = gk + 4 * id;
        double complex f[3] = {g[id], g[M + id], g[2 * M + id]};
        /* couplet grids: XX XY XZ YZ YY YX ZX ZY -> full matrix */
        double complex cXX = g[3 * M + id], cXY = g[4 * M + id], cXZ = g[5 * M + id], cYZ = g[6 * M + id];
        double complex cYY = g[7 * M + id], cYX = g[8 * M + id], cZX = g[9 * M + id], cZY = g[10 * M + id];
        double complex Cm[3][3] = {{cXX, cXY, cXZ}, {cYX, cYY, cYZ}, {cZX, cZY, -(cXX + cYY)}};
        double k2 = tk[0] * tk[0] + tk[1] * tk[1] + tk[2] * tk[2];
        double kk = sqrt(k2);
        double complex out_u[3] = {0, 0, 0}, out_d[3][3] = {{0}};
        if (id != 0) {
            double kv[3] = {tk[0], tk[1], tk[2]};
            double complex kdF = (kv[0] * f[0] + kv[1] * f[1] + kv[2] * f[2]) / k2;
            double complex Cdk[3];
            for (int a = 0; a < 3; ++a) Cdk[a] = Cm[a][0] * kv[0] + Cm[a][1] * kv[1] + Cm[a][2] * kv[2];
            double complex kCk = (Cdk[0] * kv[0] + Cdk[1] * kv[1] + Cdk[2] * kv[2]) / k2;
            double sk = sin(kk), ck = cos(kk);
            double j0 = sk / kk, j1t = 3.0 * (sk - kk * ck) / (k2 * kk);
            double Buf = tk[3] * j0 * j0;
            double Buc = -tk[3] * j0 * j1t;
            double Bdf = tk[3] * j0 * j1t;
            double Bdc = tk[3] * 9.0 * ((sk - kk * ck) / (k2 * kk)) * ((sk - kk * ck) / (k2 * kk));
            for (int a = 0; a < 3; ++a) {
                out_u[a] = (f[a] - kv[a] * kdF) * Buf;
                /* UC: velocity += i * (C.k - k (kCk)) * Buc  (Mobility.cu:417-427) */
                out_u[a] += I * (Cdk[a] - kv[a] * kCk) * Buc;
            }
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) {
                    /* DF: i (F_a k_c - k_a k_c kdF) Bdf ; DC: (Cdk_a k_c - k_a k_c kCk) Bdc */
                    out_d[a][c] = I * (f[a] * kv[c] - kv[a] * kv[c] * kdF) * Bdf
                                  + (Cdk[a] * kv[c] - kv[a] * kv[c] * kCk) * Bdc;
                }
        }
        g[id] = out_u[0]; g[M + id] = out_u[1]; g[2 * M + id] = out_u[2];
        g[3 * M + id] = out_d[0][0]; g[4 * M + id] = out_d[0][1]; g[5 * M + id] = out_d[0][2];
        g[6 * M + id] = out_d[1][2]; g[7 * M + id] = out_d[1][1]; g[8 * M + id] = out_d[1][0];
        g[9 * M + id] = out_d[2][0]; g[10 * M + id] = out_d[2][1];
    }
    for (int c = 0; c < 11; ++c) fft3(g + c * M, Nx, Ny, Nz, +1);
    contract(N, pos4, b, p, org, g, u3, D8);
    free(org);
    free(g);
}

/* contraction of the 11 real parts (Mobility_WaveSpace_ContractU/D, Mobility.cu:503-841) */
static void contract(int N, const float *pos4, const orc_box *b, const orc_pse *p, const int *org,
                     const double complex *g, double *u3, double *D8)
{
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz, P = p->P;
    const long M = (long)Nx * Ny * Nz;
    const double xisq = (double)p->xi * p->xi;
    const double prefac = pow(2.0 * xisq / M_PI / p->eta, 1.5);
    const double expfac = 2.0 * xisq / p->eta;
    const double quadW = (double)p->gridh[0] * p->gridh[1] * p->gridh[2];
    const double Lx = b->Lx, Ly = b->Ly, Lz = b->Lz;
#pragma omp parallel for
    for (int i = 0; i < N; ++i) {
        double uu[3] = {0, 0, 0}, dd[8] = {0};
        for (int tx = 0; tx < P; ++tx)
            for (int ty = 0; ty < P; ++ty)
                for (int tz = 0; tz < P; ++tz) {
                    int xi_ = org[3 * i] + tx, yi_ = org[3 * i + 1] + ty, zi_ = org[3 * i + 2] + tz;
                    xi_ = (xi_ < 0) ? xi_ + Nx : (xi_ > Nx - 1 ? xi_ - Nx : xi_);
                    yi_ = (yi_ < 0) ? yi_ + Ny : (yi_ > Ny - 1 ? yi_ - Ny : yi_);
                    zi_ = (zi_ < 0) ? zi_ + Nz : (zi_ > Nz - 1 ? zi_ - Nz : zi_);
                    double gy = (double)p->gridh[1] * yi_ - Ly / 2;
                    double gx = (double)p->gridh[0] * xi_ - Lx / 2 + (double)b->xy * gy;
                    double gz = (double)p->gridh[2] * zi_ - Lz / 2;
                    float d[3] = {(float)(gx - pos4[4 * i]), (float)(gy - pos4[4 * i + 1]), (float)(gz - pos4[4 * i + 2])};
                    float d0[3] = {d[0], d[1], d[2]};
                    orc_min_image(b, d);
                    double dx = (gx - pos4[4 * i]) + (double)(d[0] - d0[0]);
                    double dy = (gy - pos4[4 * i + 1]) + (double)(d[1] - d0[1]);
                    double dz = (gz - pos4[4 * i + 2]) + (double)(d[2] - d0[2]);
                    double w = quadW * prefac * exp(-expfac * (dx * dx + dy * dy + dz * dz));
                    long id = ((long)xi_ * Ny + yi_) * Nz + zi_;
                    for (int a = 0; a < 3; ++a) uu[a] += w * creal(g[a * M + id]);
                    for (int c = 0; c < 8; ++c) dd[c] += w * creal(g[(3 + c) * M + id]);
                }
        u3[3 * i] = uu[0]; u3[3 * i + 1] = uu[1]; u3[3 * i + 2] = uu[2];
        /* written transposed (Mobility.cu:832-833): grids XX XY XZ YZ YY YX ZX ZY -> delu */
        D8[8 * i + 0] = dd[0]; D8[8 * i + 1] = dd[5]; D8[8 * i + 2] = dd[6]; D8[8 * i + 3] = dd[7];
        D8[8 * i + 4] = dd[4]; D8[8 * i + 5] = dd[1]; D8[8 * i + 6] = dd[2]; D8[8 * i + 7] = dd[3];
    }
}

/* Far-field Brownian slip, wave-space part (Brownian_FarField.cu:113-395, 940-1011):
 * Grid1of2 places Gaussian noise on a Hermitian force grid (variance scaled by
 * T / (dt quadW)), Grid2of2 applies the square root of the wave-space Green's
 * function, then the 11 grids are inverse-transformed and contracted.  The
 * reference writes the grid from all threads at once and races where a point
 * and its conjugate both belong to the written half (the Nyquist planes);
 * here the writes happen in ascending index order, so the later point wins and
 * the grid is exactly Hermitian.  noise: 6M floats read at [q, q+6) with
 * q = idx (parity: the reference's overlapping reads) or 6 idx. */
void orc_bro_ff_wave(int N, const float *pos4, const orc_box *b, const orc_pse *p, const double *gk,
                     const float *noise, int parity, double T, double dt, double *gU)
{
    const int Nx = p->Nx, Ny = p->Ny, Nz = p->Nz;
    const long M = (long)Nx * Ny * Nz;
    const double quadW = (double)p->gridh[0] * p->gridh[1] * p->gridh[2];
    const double fac = sqrt(T / dt / quadW);
    double complex *g = calloc((size_t)11 * M, sizeof(double complex));
    for (long idx = 0; idx < M; ++idx) {
        const int kk = (int)(idx % Nz), jj = (int)((idx / Nz) % Ny), ii = (int)(idx / ((long)Ny * Nz));
        if (2 * kk >= Nz + 1) continue;
        if (kk == 0 && 2 * jj >= Ny + 1) continue;
        if (kk == 0 && jj == 0 && 2 * ii >= Nx + 1) continue;
        if (kk == 0 && jj == 0 && ii == 0) continue;
        const long q = parity ? idx : 6 * idx;
        const double re[3] = {fac * noise[q], fac * noise[q + 1], fac * noise[q + 2]};
        const double im[3] = {fac * noise[q + 3], fac * noise[q + 4], fac * noise[q + 5]};
        const int in = (ii == Nx / 2) && (Nx % 2 == 0), jn = (jj == Ny / 2) && (Ny % 2 == 0),
                  kn = (kk == Nz / 2) && (Nz % 2 == 0);
        const int ic = (ii == 0 || in) ? ii : Nx - ii, jc = (jj == 0 || jn) ? jj : Ny - jj,
                  kc = (kk == 0 || kn) ? kk : Nz - kk;
        const long cidx = ((long)ic * Ny + jc) * Nz + kc;
        const int special = (ii == 0 || in) && (jj == 0 || jn) && (kk == 0 || kn);
        for (int a = 0; a < 3; ++a) {
            if (special) {
                g[a * M + idx] = sqrt(2.0) * re[a];
            } else {
                g[a * M + idx] = re[a] + I * im[a];
                g[a * M + cidx] = re[a] - I * im[a];
            }
        }
    }
#pragma omp parallel for
    for (long idx = 0; idx < M; ++idx) {
        const double *tk = gk + 4 * idx;
        const double ksq = tk[0] * tk[0] + tk[1] * tk[1] + tk[2] * tk[2];
        const double k = sqrt(ksq);
        const double B = (idx == 0) ? 0.0 : sqrt(tk[3]);
        const double SU = (idx == 0) ? 0.0 : sin(k) / k;
        const double SD = (idx == 0) ? 0.0 : 3.0 * (sin(k) - k * cos(k)) / (ksq * k);
        double complex f[3] = {g[idx], g[M + idx], g[2 * M + idx]};
        double complex kdF = (idx == 0) ? 0.0 : (tk[0] * f[0] + tk[1] * f[1] + tk[2] * f[2]) / ksq;
        double complex W[3];
        for (int a = 0; a < 3; ++a) W[a] = (f[a] - tk[a] * kdF) * B;
        for (int a = 0; a < 3; ++a) g[a * M + idx] = SU * W[a];
        /* gradient grids in the order XX XY XZ YZ YY YX ZX ZY: component (a, c) = i SD W_a k_c */
        static const int sa[8] = {0, 0, 0, 1, 1, 1, 2, 2}, sc[8] = {0, 1, 2, 2, 1, 0, 0, 1};
        for (int s2 = 0; s2 < 8; ++s2) g[(3 + s2) * M + idx] = I * SD * W[sa[s2]] * tk[sc[s2]];
    }
    for (int c = 0; c < 11; ++c) fft3(g + c * M, Nx, Ny, Nz, +1);
    int *org = stencil_origins(N, pos4, b, p);
    double *u3 = calloc(3 * (size_t)N, sizeof(double)), *D8 = calloc(8 * (size_t)N, sizeof(double));
    contract(N, pos4, b, p, org, g, u3, D8);
    for (int i = 0; i < N; ++i) {
        double Wv[3], E5[5];
        d2we(D8 + 8 * i, Wv, E5);
        for (int a = 0; a < 3; ++a) { gU[6 * i + a] = u3[3 * i + a]; gU[6 * i + 3 + a] = Wv[a]; }
        for (int c = 0; c < 5; ++c) gU[6 * (size_t)N + 5 * i + c] = E5[c];
    }
    free(org); free(u3); free(D8); free(g);
}

/* Mobility_GeneralizedMobility (Mobility.cu:1488-1542) */
void orc_mobility(int N, const float *pos4, const orc_box *b, const orc_pse *p, const float *tab,
                  const double *gk, const int *nneigh, const int *head, const int *nbrs,
                  const double *gF, double *gU, int parts)
{
    double *F3 = malloc(sizeof(double) * 3 * N), *C8 = malloc(sizeof(double) * 8 * N);
    double *u3 = calloc(3 * (size_t)N, sizeof(double)), *D8 = calloc(8 * (size_t)N, sizeof(double));
    for (int i = 0; i < N; ++i) {
        for (int a = 0; a < 3; ++a) F3[3 * i + a] = gF[6 * i + a];
        ts2c(gF + 6 * i + 3, gF + 6 * (size_t)N + 5 * i, C8 + 8 * i);
    }
    if (parts & 2) wave_space(N, pos4, b, p, gk, F3, C8, u3, D8);
    if (parts & 1) {
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < N; ++i) {
            double u[3], D[8];
            real_space_i(i, pos4, b, p, tab, nneigh, head, nbrs, F3, C8, u, D);
            for (int a = 0; a < 3; ++a) u3[3 * i + a] += u[a];
            for (int c = 0; c < 8; ++c) D8[8 * i + c] += D[c];
        }
    }
    for (int i = 0; i < N; ++i) {
        double W[3], E5[5];
        d2we(D8 + 8 * i, W, E5);
        for (int a = 0; a < 3; ++a) { gU[6 * i + a] = u3[3 * i + a]; gU[6 * i + 3 + a] = W[a]; }
        for (int c = 0; c < 5; ++c) gU[6 * (size_t)N + 5 * i + c] = E5[c];
    }
    free(F3); free(C8); free(u3); free(D8);
}

/* ------------------------------------------------------- preconditioner */
/* Precondition_GetPrunedNneigh/GetPrunedNlist (Precondition.cu:70-242) */
int orc_prune(int N, const float *pos4, const orc_box *b, const int *nneigh, const int *head,
              const int *nbrs, float rp, int *pn, int *phead, int *plist, int *pless, int cap)
{
    const float rp2 = rp * rp;
    int tot = 0;
    for (int i = 0; i < N; ++i) {
        phead[i] = tot;
        int start = tot;
        if (tot < cap) plist[tot] = i;
        ++tot;
        for (int n = 0; n < nneigh[i]; ++n) {
            int j = nbrs[head[i] + n];
            float d[3] = {pos4[4 * i] - pos4[4 * j], pos4[4 * i + 1] - pos4[4 * j + 1], pos4[4 * i + 2] - pos4[4 * j + 2]};
            orc_min_image(b, d);
            float d2 = d[0] * d[0]; float t = d[1] * d[1]; d2 = d2 + t; t = d[2] * d[2]; d2 = d2 + t;
            if (d2 < rp2) { if (tot < cap) plist[tot] = j; ++tot; }
        }
        pn[i] = tot - start;
        if (tot <= cap) {
            /* insertion sort ascending */
            int *A = plist + start;
            for (int a = 1; a < pn[i]; ++a) {
                int key = A[a], c = a - 1;
                while (c >= 0 && A[c] > key) { A[c + 1] = A[c]; --c; }
                A[c + 1] = key;
            }
            int nl = 0;
            for (int a = 0; a < pn[i]; ++a) nl += (A[a] < i);
            pless[i] = nl;
        }
    }
    phead[N] = tot;
    return tot;
}

/* Precondition_PreProcess + Precondition_RFU_kernel + coo2csr
 * (Precondition.cu:443-501, 530-924).  Rows in COO order are ascending, so
 * the CSR is the COO with a row pointer. */
int orc_build_rfu_csr(int N, const float *pos4, const orc_box *b, const int *pn, const int *phead,
                      const int *plist, const int *pless, const orc_lub *t, orc_csr *A)
{
    int *offset = malloc(sizeof(int) * (N + 1));
    int *nepp = malloc(sizeof(int) * N);
    long nnz = 0;
    for (int i = 0; i < N; ++i) {
        nepp[i] = (pn[i] > 1) ? 18 * pn[i] : 3;
        offset[i] = (int)nnz;
        nnz += 2 * nepp[i];
    }
    offset[N] = (int)nnz;
    A->n = 6 * N; A->nnz = (int)nnz;
    A->rowptr = malloc(sizeof(int) * (6 * N + 1));
    A->col = malloc(sizeof(int) * nnz);
    A->val = calloc(nnz, sizeof(double));
    int *row = malloc(sizeof(int) * nnz);
    const float rp2 = t->rp * t->rp;
    for (int i = 0; i < N; ++i) {
        const int nn = pn[i], off = offset[i], obc = nepp[i];
        int flag = 0;
        for (int s = 0; s < nn; ++s) {
            int j = plist[phead[i] + s];
            float d[3];
            float d2 = pair_d2(b, pos4 + 4 * i, pos4 + 4 * j, d);
            if (!(d2 < rp2 && j != i)) continue;
            float dist = sqrtf(d2);
            double v[10];
            if ((double)dist <= 2.0 + (double)t->table_min) {
                for (int c = 0; c < 10; ++c) v[c] = t->vals[c];          /* row 0 clamp */
            } else {
                lub_lookup(t, dist, 0, 10, -1, v);
            }
            double XA11 = v[0], XA12 = v[1], YA11 = v[2], YA12 = v[3], YB11 = v[4], YB12 = v[5];
            double XC11 = v[6], XC12 = v[7], YC11 = v[8], YC12 = v[9];
            double r[3] = {d[0] / (double)dist, d[1] / (double)dist, d[2] / (double)dist};
            double er[3][3] = {{0, r[2], -r[1]}, {-r[2], 0, r[0]}, {r[1], -r[0], 0}};
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) {
                    double rr = r[a] * r[c], imrr = (a == c ? 1.0 : 0.0) - rr;
                    double An = XA12 * rr + YA12 * imrr, As = XA11 * rr + YA11 * imrr;
                    double Bn = YB12 * er[a][c], Bs = YB11 * er[a][c];
                    double Cn = XC12 * rr + YC12 * imrr, Cs = XC11 * rr + YC11 * imrr;
                    int rw = 6 * i + a;
                    int on = off + 6 * nn * a + 6 * s + c;
                    int os = off + 6 * nn * a + 6 * pless[i] + c;
                    row[on] = rw;           A->col[on] = 6 * j + c;           A->val[on] = An;
                    row[on + 3] = rw;       A->col[on + 3] = 6 * j + 3 + c;   A->val[on + 3] = Bn;
                    row[obc + on] = rw + 3; A->col[obc + on] = 6 * j + c;     A->val[obc + on] = Bn;
                    row[obc + on + 3] = rw + 3; A->col[obc + on + 3] = 6 * j + 3 + c; A->val[obc + on + 3] = Cn;
                    if (flag < 9) {
                        row[os] = rw; A->col[os] = 6 * i + c;
                        row[os + 3] = rw; A->col[os + 3] = 6 * i + 3 + c;
                        row[obc + os] = rw + 3; A->col[obc + os] = 6 * i + c;
                        row[obc + os + 3] = rw + 3; A->col[obc + os + 3] = 6 * i + 3 + c;
                    }
                    A->val[os] += As; A->val[os + 3] += -Bs; A->val[obc + os] += Bs; A->val[obc + os + 3] += Cs;
                    ++flag;
                }
        }
        if (flag == 0) {
            for (int a = 0; a < 3; ++a) {
                row[off + a] = 6 * i + a; A->col[off + a] = 6 * i + a; A->val[off + a] = 0.0;
                row[obc + off + a] = 6 * i + 3 + a; A->col[obc + off + a] = 6 * i + 3 + a; A->val[obc + off + a] = 0.0;
            }
        }
    }
    /* coo2csr */
    for (int r = 0; r <= 6 * N; ++r) A->rowptr[r] = 0;
    for (long e = 0; e < nnz; ++e) A->rowptr[row[e] + 1]++;
    for (int r = 0; r < 6 * N; ++r) A->rowptr[r + 1] += A->rowptr[r];
    free(row); free(offset); free(nepp);
    return 0;
}

/* -------- reverse Cuthill-McKee (George & Liu SPARSPAK GENRCM; rcm.cpp:870-3574) -------- */
/* BFS level structure from root over unmasked nodes; returns number of levels */
static int rcm_levels(int root, const int *xadj, const int *adj, int *mask, int *lvl_start, int *ls, int *size)
{
    mask[root] = 0;
    ls[0] = root;
    int nlev = 0, lend = 0, cnt = 1;
    for (;;) {
        int lbeg = lend;
        lend = cnt;
        lvl_start[nlev++] = lbeg;
        for (int q = lbeg; q < lend; ++q) {
            int node = ls[q];
            for (int e = xadj[node]; e < xadj[node + 1]; ++e) {
                int nb = adj[e];
                if (mask[nb]) { ls[cnt++] = nb; mask[nb] = 0; }
            }
        }
        if (cnt - lend <= 0) break;
    }
    lvl_start[nlev] = lend;
    for (int q = 0; q < cnt; ++q) mask[ls[q]] = 1;
    *size = cnt;
    return nlev;
}

/* pseudo-peripheral root (root_find) */
static int rcm_root(int root, const int *xadj, const int *adj, int *mask, int *lvl_start, int *ls)
{
    int size;
    int nlev = rcm_levels(root, xadj, adj, mask, lvl_start, ls, &size);
    if (nlev == 1 || nlev == size) return root;
    for (;;) {
        int mindeg = size;
        int jst = lvl_start[nlev - 1];
        root = ls[jst];
        if (jst < size - 1) {
            for (int q = jst; q < size; ++q) {
                int node = ls[q], nd = 0;
                for (int e = xadj[node]; e < xadj[node + 1]; ++e) nd += (mask[adj[e]] > 0);
                if (nd < mindeg) { root = node; mindeg = nd; }
            }
        }
        int nlev2 = rcm_levels(root, xadj, adj, mask, lvl_start, ls, &size);
        if (nlev2 <= nlev) break;
        nlev = nlev2;
        if (size <= nlev) break;
    }
    return root;
}

/* Cuthill-McKee from root (degrees over the unmasked component), then reversed */
static int rcm_component(int root, const int *xadj, const int *adj, int *mask, int *perm, int *deg, int *vis)
{
    /* degrees within the component (self loops counted, as the reference) */
    int cnt = 1;
    perm[0] = root; vis[root] = 1;
    for (int q = 0; q < cnt; ++q) {
        int node = perm[q], d = 0;
        for (int e = xadj[node]; e < xadj[node + 1]; ++e) {
            int nb = adj[e];
            if (mask[nb]) { ++d; if (!vis[nb]) { vis[nb] = 1; perm[cnt++] = nb; } }
        }
        deg[node] = d;
    }
    const int size = cnt;
    for (int q = 0; q < size; ++q) vis[perm[q]] = 0;
    mask[root] = 0;
    if (size == 1) return 1;
    int lend = 0, lnbr = 1;
    while (lend < lnbr) {
        int lbeg = lend;
        lend = lnbr;
        for (int q = lbeg; q < lend; ++q) {
            int node = perm[q];
            int fnbr = lnbr;
            for (int e = xadj[node]; e < xadj[node + 1]; ++e) {
                int nb = adj[e];
                if (mask[nb]) { mask[nb] = 0; perm[lnbr++] = nb; }
            }
            if (lnbr - fnbr <= 1) continue;
            /* insertion sort by increasing degree as rcm.cpp:3340-3362 runs it: its
             * loop `while ( fnbr < l )` never compares the first new neighbour, so
             * perm[fnbr] stays first and only perm[fnbr+1..lnbr) is sorted (stable) */
            for (int k = fnbr + 1; k < lnbr; ++k) {
                int nb = perm[k], l = k - 1;
                while (l > fnbr && deg[perm[l]] > deg[nb]) { perm[l + 1] = perm[l]; --l; }
                perm[l + 1] = nb;
            }
        }
    }
    for (int a = 0, c = size - 1; a < c; ++a, --c) { int tmp = perm[a]; perm[a] = perm[c]; perm[c] = tmp; }
    return size;
}

void orc_rcm(int N, const int *phead, const int *plist, int *perm)
{
    int *mask = malloc(sizeof(int) * N), *lvl = malloc(sizeof(int) * (N + 1));
    int *deg = malloc(sizeof(int) * N), *vis = calloc(N, sizeof(int));
    for (int i = 0; i < N; ++i) mask[i] = 1;
    int num = 0;
    for (int i = 0; i < N; ++i) {
        if (!mask[i]) continue;
        int root = rcm_root(i, phead, plist, mask, lvl, perm + num);
        num += rcm_component(root, phead, plist, mask, perm + num, deg, vis);
        if (num >= N) break;
    }
    free(mask); free(lvl); free(deg); free(vis);
}

static int cmp_int_pair(const void *a, const void *b)
{
    const int *x = a, *y = b;
    return (x[0] > y[0]) - (x[0] < y[0]);
}

/* csrperm: B = A(p, p), p[6i+c] = 6 perm[i] + c; columns sorted */
int orc_permute(const orc_csr *A, const int *perm, int N, orc_csr *B)
{
    const int n = 6 * N;
    int *p = malloc(sizeof(int) * n), *ip = malloc(sizeof(int) * n);
    for (int i = 0; i < N; ++i)
        for (int c = 0; c < 6; ++c) p[6 * i + c] = 6 * perm[i] + c;
    for (int i = 0; i < n; ++i) ip[p[i]] = i;
    B->n = n; B->nnz = A->nnz;
    B->rowptr = malloc(sizeof(int) * (n + 1));
    B->col = malloc(sizeof(int) * A->nnz);
    B->val = malloc(sizeof(double) * A->nnz);
    B->rowptr[0] = 0;
    int *pair = malloc(sizeof(int) * 2 * 64);
    int cappair = 64;
    for (int i = 0; i < n; ++i) {
        int r = p[i], len = A->rowptr[r + 1] - A->rowptr[r];
        if (len > cappair) { cappair = len; pair = realloc(pair, sizeof(int) * 2 * cappair); }
        for (int e = 0; e < len; ++e) { pair[2 * e] = ip[A->col[A->rowptr[r] + e]]; pair[2 * e + 1] = A->rowptr[r] + e; }
        qsort(pair, len, 2 * sizeof(int), cmp_int_pair);
        B->rowptr[i + 1] = B->rowptr[i] + len;
        for (int e = 0; e < len; ++e) {
            B->col[B->rowptr[i] + e] = pair[2 * e];
            B->val[B->rowptr[i] + e] = A->val[pair[2 * e + 1]];
        }
    }
    free(pair); free(p); free(ip);
    return 0;
}

/* IC(0) on the lower triangle of the stored pattern (csric02 semantics,
 * Precondition.cu:1116-1307), after adding lambda*(1 | 4/3) to the diagonal
 * (Precondition_AddIdentity_kernel).  Upper entries are zeroed afterwards. */
int orc_ichol(orc_csr *B, double lambda)
{
    const int n = B->n;
    int *diag = malloc(sizeof(int) * n);
    for (int i = 0; i < n; ++i) {
        diag[i] = -1;
        for (int e = B->rowptr[i]; e < B->rowptr[i + 1]; ++e)
            if (B->col[e] == i) { diag[i] = e; break; }
        if (diag[i] < 0) { free(diag); return -1; }
        B->val[diag[i]] += lambda * (((i % 6) < 3) ? 1.0 : 1.33333333);
    }
    int fail = 0;
    for (int i = 0; i < n && !fail; ++i) {
        /* row i: for each j < i in pattern: L_ij = (A_ij - sum_k<j L_ik L_jk) / L_jj */
        for (int e = B->rowptr[i]; e < B->rowptr[i + 1]; ++e) {
            int j = B->col[e];
            if (j > i) break;
            double s = B->val[e];
            /* sparse dot of row i and row j over columns k < j */
            int a = B->rowptr[i], bb = B->rowptr[j];
            while (a < e && bb < B->rowptr[j + 1]) {
                int ca = B->col[a], cb = B->col[bb];
                if (cb >= j) break;
                if (ca == cb) { s -= B->val[a] * B->val[bb]; ++a; ++bb; }
                else if (ca < cb) ++a;
                else ++bb;
            }
            if (j < i) {
                B->val[e] = s / B->val[diag[j]];
            } else {
                if (!(s > 0.0)) { fail = 1; break; }
                B->val[e] = sqrt(s);
            }
        }
    }
    for (int i = 0; i < n; ++i)
        for (int e = B->rowptr[i]; e < B->rowptr[i + 1]; ++e)
            if (B->col[e] > i) B->val[e] = 0.0;
    free(diag);
    return fail;
}

/* Precondition_Saddle_RFUmultiply (Precondition.cu:1802-1890): y = P^T L^-T L^-1 P x */
void orc_precond_apply(const orc_csr *L, const int *perm, int N, const double *x, double *y)
{
    const int n = 6 * N;
    double *t = malloc(sizeof(double) * n);
    for (int i = 0; i < N; ++i)
        for (int c = 0; c < 6; ++c) t[6 * i + c] = x[6 * perm[i] + c];
    /* forward: L z = t */
    for (int i = 0; i < n; ++i) {
        double s = t[i], dg = 1.0;
        for (int e = L->rowptr[i]; e < L->rowptr[i + 1]; ++e) {
            int j = L->col[e];
            if (j < i) s -= L->val[e] * t[j];
            else if (j == i) dg = L->val[e];
        }
        t[i] = s / dg;
    }
    /* backward: L^T w = z (column sweep) */
    for (int i = n - 1; i >= 0; --i) {
        double dg = 1.0;
        for (int e = L->rowptr[i]; e < L->rowptr[i + 1]; ++e) if (L->col[e] == i) dg = L->val[e];
        t[i] /= dg;
        for (int e = L->rowptr[i]; e < L->rowptr[i + 1]; ++e) {
            int j = L->col[e];
            if (j < i) t[j] -= L->val[e] * t[i];
        }
    }
    for (int i = 0; i < N; ++i)
        for (int c = 0; c < 6; ++c) y[6 * perm[i] + c] = t[6 * i + c];
    free(t);
}

void orc_csr_free(orc_csr *A)
{
    free(A->rowptr); free(A->col); free(A->val);
    A->rowptr = NULL; A->col = NULL; A->val = NULL;
}

/* ------------------------------------------------------------ saddle */
/* Saddle_Multiply (Saddle.cu:80-151) */
void orc_saddle_multiply(const orc_system *s, const double *x, double *y)
{
    const int N = s->N;
    orc_mobility(N, s->pos4, &s->box, s->pse, s->ewald_tab, s->gridk4, s->nneigh, s->head, s->nbrs, x, y, 3);
    for (int k = 0; k < 6 * N; ++k) y[k] += x[11 * N + k];
    double *r = malloc(sizeof(double) * 6 * N);
    orc_rfu(N, s->pos4, &s->box, s->nneigh, s->head, s->nbrs, s->lub, x + 11 * N, r);
    for (int k = 0; k < 6 * N; ++k) y[11 * N + k] = x[k] - r[k];
    free(r);
}

/* Saddle_Preconditioner (Saddle.cu:171-264) */
void orc_saddle_precond(const orc_system *s, const double *b, double *x)
{
    const int N = s->N;
    double *gF = malloc(sizeof(double) * 6 * N), *gU = malloc(sizeof(double) * 6 * N);
    orc_precond_apply(s->L, s->perm, N, b, gF);
    orc_precond_apply(s->L, s->perm, N, b + 11 * N, gU);
    for (int k = 0; k < 6 * N; ++k) x[k] = b[k] - gF[k] + gU[k];
    for (int k = 6 * N; k < 11 * N; ++k) x[k] = b[k];
    for (int k = 0; k < 6 * N; ++k) x[11 * N + k] = gF[k] - gU[k];
    free(gF); free(gU);
}

static double nrm2(const double *a, long n) { double s = 0; for (long k = 0; k < n; ++k) s += a[k] * a[k]; return sqrt(s); }

/* left-preconditioned restarted GMRES with CUSP's stopping rule
 * (Solvers.cu:55-126): stop when |s_{i+1}| <= tol * ||b|| */
int orc_gmres(const orc_system *s, const double *b, double *x, int R, int maxit, double tol,
              int use_precond, double *resid_out)
{
    const long n = 17L * s->N;
    double *V = malloc(sizeof(double) * n * (R + 1));
    double *w = malloc(sizeof(double) * n), *t = malloc(sizeof(double) * n);
    double *H = calloc((size_t)(R + 1) * R, sizeof(double));
    double *cs = malloc(sizeof(double) * R), *sn = malloc(sizeof(double) * R), *g = malloc(sizeof(double) * (R + 1));
    const double bn = nrm2(b, n), target = tol * bn;
    int iters = 0;
    double resid = 0;
    for (;;) {
        orc_saddle_multiply(s, x, w);
        for (long k = 0; k < n; ++k) w[k] -= b[k];
        if (use_precond) { orc_saddle_precond(s, w, t); memcpy(w, t, sizeof(double) * n); }
        double beta = nrm2(w, n);
        resid = beta;
        if (resid <= target || iters >= maxit || beta == 0.0) break;
        for (long k = 0; k < n; ++k) V[k] = -w[k] / beta;
        memset(g, 0, sizeof(double) * (R + 1));
        g[0] = beta;
        int i = -1;
        do {
            ++i; ++iters;
            orc_saddle_multiply(s, V + (long)i * n, t);
            if (use_precond) orc_saddle_precond(s, t, w); else memcpy(w, t, sizeof(double) * n);
            for (int k = 0; k <= i; ++k) {
                double h = 0;
                const double *vk = V + (long)k * n;
                for (long q = 0; q < n; ++q) h += w[q] * vk[q];
                H[k * R + i] = h;
                for (long q = 0; q < n; ++q) w[q] -= h * vk[q];
            }
            double hn = nrm2(w, n);
            H[(i + 1) * R + i] = hn;
            for (long q = 0; q < n; ++q) V[(long)(i + 1) * n + q] = (hn != 0.0) ? w[q] / hn : 0.0;
            for (int k = 0; k < i; ++k) {
                double a = H[k * R + i], c = H[(k + 1) * R + i];
                H[k * R + i] = cs[k] * a + sn[k] * c;
                H[(k + 1) * R + i] = -sn[k] * a + cs[k] * c;
            }
            double a = H[i * R + i], c = H[(i + 1) * R + i];
            if (c == 0.0) { cs[i] = 1.0; sn[i] = 0.0; }
            else if (fabs(c) > fabs(a)) { double tt = a / c; sn[i] = 1.0 / sqrt(1.0 + tt * tt); cs[i] = tt * sn[i]; }
            else { double tt = c / a; cs[i] = 1.0 / sqrt(1.0 + tt * tt); sn[i] = tt * cs[i]; }
            H[i * R + i] = cs[i] * a + sn[i] * c;
            H[(i + 1) * R + i] = 0.0;
            g[i + 1] = -sn[i] * g[i];
            g[i] = cs[i] * g[i];
            resid = fabs(g[i + 1]);
        } while (!(resid <= target) && iters < maxit && i < R - 1);
        /* back substitution and update */
        for (int j = i; j >= 0; --j) {
            g[j] /= H[j * R + j];
            for (int k = j - 1; k >= 0; --k) g[k] -= H[k * R + j] * g[j];
        }
        for (int j = 0; j <= i; ++j)
            for (long q = 0; q < n; ++q) x[q] += g[j] * V[(long)j * n + q];
        if (resid <= target || iters >= maxit) break;
    }
    if (resid_out) *resid_out = resid / (bn > 0 ? bn : 1.0);
    free(V); free(w); free(t); free(H); free(cs); free(sn); free(g);
    return iters;
}

/* ---------------------------------------------------- forces / Euler */
/* Stokes_SetForce_manually_kernel (Helper_Stokes.cu:65-172) */
void orc_force_law(int N, const float *pos4, const orc_box *b, const int *nneigh, const int *head,
                   const int *nbrs, double ndsr, double k_n, double kappa, double beta, double epsq,
                   double T_ext, double *F6)
{
    const double h_rough = sqrtf((float)epsq), rcol = (float)(2.0f + 1.0f * (float)h_rough), F0 = 1.0 / ndsr, Ham = F0 * beta;
    for (int i = 0; i < N; ++i) {
        double F[3] = {0, 0, 0};
        for (int n = 0; n < nneigh[i]; ++n) {
            int j = nbrs[head[i] + n];
            float d[3];
            float d2 = pair_d2(b, pos4 + 4 * i, pos4 + 4 * j, d);
            float dist = sqrtf(d2);
            /* the gap is a float difference in the reference (Helper_Stokes.cu:127-128);
             * with k_n = 1e4 its rounding dominates the collision force, so keep it */
            double gap = (double)(dist - (float)rcol);
            double mag = 0.0;
            if (gap >= 0. && (double)dist <= 2.0 + 10.0 / kappa)
                mag = Ham / (12. * (gap * gap + epsq)) - F0 * exp(-kappa * gap);
            if (gap < 0.) mag = Ham / (12. * epsq) - F0 - k_n * fabs(gap);
            for (int a = 0; a < 3; ++a) F[a] += mag * d[a] / (double)dist;
        }
        for (int a = 0; a < 3; ++a) F6[6 * i + a] = F[a];
        F6[6 * i + 3] = 0.0; F6[6 * i + 4] = 0.0; F6[6 * i + 5] = T_ext;
    }
}

/* Integrator_ExplicitEuler_Shear_kernel (Integrator.cu:126-228) */
void orc_euler_shear(int N, float *pos4, float *ori3, float *pos_gb4, int *image3, const orc_box *b,
                     const double *vel6, double B1, const float *mask1, const float *noise3,
                     double dt, double shear_rate)
{
    for (int i = 0; i < N; ++i) {
        double p[3] = {pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2]};
        double v[3] = {vel6[6 * i], vel6[6 * i + 1], vel6[6 * i + 2]};
        double w[3] = {vel6[6 * i + 3], vel6[6 * i + 4], vel6[6 * i + 5]};
        v[0] += shear_rate * p[1];
        w[2] -= shear_rate / 2.;
        if (noise3) for (int a = 0; a < 3; ++a) w[a] += noise3[3 * i + a];
        double q0 = 1.0, qv[3] = {0, 0, 0};
        double aw = sqrt(dot3(w, w));
        if (aw > 1e-6) {
            double ang = aw * dt;
            q0 = cos(ang / 2.);
            for (int a = 0; a < 3; ++a) qv[a] = w[a] / aw * sin(ang / 2.);
        }
        double pv[3] = {ori3[3 * i], ori3[3 * i + 1], ori3[3 * i + 2]};
        double m1 = mask1 ? mask1[i] : 1.0;
        for (int a = 0; a < 3; ++a) v[a] -= (-2. / 3. * B1 * m1) * pv[a];
        pos_gb4[4 * i] += (float)((v[0] - shear_rate * p[1] + shear_rate * pos_gb4[4 * i + 1]) * dt);
        pos_gb4[4 * i + 1] += (float)(v[1] * dt);
        pos_gb4[4 * i + 2] += (float)(v[2] * dt);
        float np[3] = {(float)(p[0] + v[0] * dt), (float)(p[1] + v[1] * dt), (float)(p[2] + v[2] * dt)};
        orc_wrap(b, np, image3 + 3 * i);
        pos4[4 * i] = np[0]; pos4[4 * i + 1] = np[1]; pos4[4 * i + 2] = np[2];
        double qq = dot3(qv, qv), qp = dot3(qv, pv);
        double cr[3] = {qv[1] * pv[2] - qv[2] * pv[1], qv[2] * pv[0] - qv[0] * pv[2], qv[0] * pv[1] - qv[1] * pv[0]};
        double pn[3];
        for (int a = 0; a < 3; ++a) pn[a] = (q0 * q0 - qq) * pv[a] + 2. * qp * qv[a] + 2. * q0 * cr[a];
        double m = sqrt(dot3(pn, pn));
        for (int a = 0; a < 3; ++a) ori3[3 * i + a] = (float)(pn[a] / m);
    }
}

/* director rotation pv <- q pv q* over angle |w| tau (Integrator.cu:278-286, 312-321) */
static void orc_rotate(const double w[3], double tau, double pv[3])
{
    double q0 = 1.0, qv[3] = {0, 0, 0};
    const double aw = sqrt(dot3(w, w));
    if (aw > 1e-6) {
        const double ang = aw * tau;
        q0 = cos(ang / 2.);
        for (int a = 0; a < 3; ++a) qv[a] = w[a] / aw * sin(ang / 2.);
    }
    const double qq = dot3(qv, qv), qp = dot3(qv, pv);
    const double cr[3] = {qv[1] * pv[2] - qv[2] * pv[1], qv[2] * pv[0] - qv[0] * pv[2], qv[0] * pv[1] - qv[1] * pv[0]};
    double pn[3];
    for (int a = 0; a < 3; ++a) pn[a] = (q0 * q0 - qq) * pv[a] + 2. * qp * qv[a] + 2. * q0 * cr[a];
    const double m = sqrt(dot3(pn, pn));
    for (int a = 0; a < 3; ++a) pv[a] = pn[a] / m;
}

/* Integrator_ExplicitEuler1_Shear_kernel (Integrator.cu:232-326) */
void orc_rk2_stage1(int N, const float *pos_in4, const float *ori_in3, float *pos_out4, float *ori_out3,
                    float *pos_gb4, const double *vel6, double B1, const float *mask1, const float *noise3,
                    double dt, double shear_rate)
{
    for (int i = 0; i < N; ++i) {
        double p[3] = {pos_in4[4 * i], pos_in4[4 * i + 1], pos_in4[4 * i + 2]};
        double v[3] = {vel6[6 * i], vel6[6 * i + 1], vel6[6 * i + 2]};
        double w[3] = {vel6[6 * i + 3], vel6[6 * i + 4], vel6[6 * i + 5]};
        v[0] += shear_rate * p[1];
        w[2] -= shear_rate / 2.;
        if (noise3) for (int a = 0; a < 3; ++a) w[a] += noise3[3 * i + a];
        double pv[3] = {ori_in3[3 * i], ori_in3[3 * i + 1], ori_in3[3 * i + 2]};
        const double m1 = mask1 ? mask1[i] : 1.0;
        for (int a = 0; a < 3; ++a) v[a] -= (-2. / 3. * B1 * m1) * pv[a];
        pos_gb4[4 * i] += (float)((v[0] - shear_rate * p[1] + shear_rate * pos_gb4[4 * i + 1]) * 0.5 * dt);
        pos_gb4[4 * i + 1] += (float)(v[1] * 0.5 * dt);
        pos_gb4[4 * i + 2] += (float)(v[2] * 0.5 * dt);
        for (int a = 0; a < 3; ++a) pos_out4[4 * i + a] = (float)(p[a] + v[a] * dt);
        pos_out4[4 * i + 3] = pos_in4[4 * i + 3];
        orc_rotate(w, dt, pv);
        for (int a = 0; a < 3; ++a) ori_out3[3 * i + a] = (float)pv[a];
    }
}

/* Integrator_RK_Shear_kernel (Integrator.cu:329-438) */
void orc_rk2_stage2(int N, double c1, const float *pos1, const float *ori1, double c2, const float *pos2,
                    const float *ori2, float *pos_out4, float *ori_out3, float *pos_gb4, const double *vel6,
                    double B1, const float *mask1, const float *noise3, int *image3, const orc_box *b, double c3,
                    double dt, double shear_rate)
{
    for (int i = 0; i < N; ++i) {
        double p[3], pv[3];
        for (int a = 0; a < 3; ++a) {
            p[a] = c1 * pos1[4 * i + a] + c2 * pos2[4 * i + a];
            pv[a] = ori1[3 * i + a] * c1 + ori2[3 * i + a] * c2;
        }
        const double m0 = sqrt(dot3(pv, pv));
        for (int a = 0; a < 3; ++a) pv[a] /= m0;
        double v[3] = {vel6[6 * i], vel6[6 * i + 1], vel6[6 * i + 2]};
        double w[3] = {vel6[6 * i + 3], vel6[6 * i + 4], vel6[6 * i + 5]};
        v[0] += shear_rate * p[1];
        w[2] -= shear_rate / 2.;
        if (noise3) for (int a = 0; a < 3; ++a) w[a] += noise3[3 * i + a];
        const double m1 = mask1 ? mask1[i] : 1.0;
        for (int a = 0; a < 3; ++a) v[a] -= (-2. / 3. * B1 * m1) * pv[a];
        pos_gb4[4 * i] += (float)((v[0] - shear_rate * p[1] + shear_rate * pos_gb4[4 * i + 1]) * 0.5 * dt);
        pos_gb4[4 * i + 1] += (float)(v[1] * 0.5 * dt);
        pos_gb4[4 * i + 2] += (float)(v[2] * 0.5 * dt);
        float np[3] = {(float)(p[0] + v[0] * dt * c3), (float)(p[1] + v[1] * dt * c3), (float)(p[2] + v[2] * dt * c3)};
        orc_wrap(b, np, image3 + 3 * i);
        for (int a = 0; a < 3; ++a) pos_out4[4 * i + a] = np[a];
        pos_out4[4 * i + 3] = 0.f;
        orc_rotate(w, dt * c3, pv);
        for (int a = 0; a < 3; ++a) ori_out3[3 * i + a] = (float)pv[a];
    }
}

int orc_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------- direct (spectrally converged) Ewald sum */
/* The FTS grand mobility of the periodic RPY suspension without the PSE
 * discretisation: the same Ewald splitting as Mobility.cu:1345-1542 (real-space
 * functions of Stokes.cc:412-666, wave-space operator of Mobility.cu:315-479), but
 * the real-space functions evaluated in FP64 at every pair and periodic image within
 * r_cut, the wave-space part summed directly over every reciprocal vector with
 * |k| <= k_cut (structure factors, no Gaussian gridding, no FFT), the exact 2 pi,
 * and self terms in double.  With erfc(xi r_cut) and exp(-k_cut^2 / 4 xi^2) below
 * ~1e-14 the result is the Ewald sum itself, independent of xi: the reference
 * against which the PSE's method error is measured. */
static void ewald_self_double(double xi, double s[2])
{
    const double pi = M_PI, sp = sqrt(pi);
    s[0] = (1. + 4. * sp * xi * erfc(2. * xi) - exp(-4. * xi * xi)) / (4. * sp * xi);
    s[1] = -3. * erfc(2. * xi) / 10. - 3. / (80. * sp * xi * xi * xi) - 9. / (40. * sp * xi)
           + 3. * exp(-4. * xi * xi) * (1. + 10. * xi * xi) / (80. * sp * xi * xi * xi);
}

void orc_mobility_direct(int N, const float *pos4, const orc_box *b, double xi, double r_cut, double k_cut,
                         const double *gF, double *gU, int parts)
{
    double *F3 = malloc(sizeof(double) * 3 * N), *C8 = malloc(sizeof(double) * 8 * N);
    double *u3 = calloc(3 * (size_t)N, sizeof(double)), *D8 = calloc(8 * (size_t)N, sizeof(double));
    for (int i = 0; i < N; ++i) {
        for (int a = 0; a < 3; ++a) F3[3 * i + a] = gF[6 * i + a];
        ts2c(gF + 6 * i + 3, gF + 6 * (size_t)N + 5 * i, C8 + 8 * i);
    }
    const double Lx = b->Lx, Ly = b->Ly, Lz = b->Lz, xy = b->xy;
    const double V = Lx * Ly * Lz;
    if (parts & 1) {
        double self[2];
        ewald_self_double(xi, self);
        /* periodic images n1 a1 + n2 a2 + n3 a3, a1 = (Lx,0,0), a2 = (xy Ly, Ly, 0), a3 = (0,0,Lz) */
        const int n3m = (int)ceil(r_cut / Lz) + 1, n2m = (int)ceil(r_cut / Ly) + 1;
        const int n1m = (int)ceil((r_cut + fabs(xy) * Ly * n2m) / Lx) + 1;
#pragma omp parallel for schedule(dynamic, 8)
        for (int i = 0; i < N; ++i) {
            const double *Fi = F3 + 3 * i, *Ci = C8 + 8 * i;
            double u[3], D[8];
            u[0] = self[0] * Fi[0]; u[1] = self[0] * Fi[1]; u[2] = self[0] * Fi[2];
            D[0] = self[1] * (Ci[0] - 4. * Ci[0]); D[1] = self[1] * (Ci[1] - 4. * Ci[5]);
            D[2] = self[1] * (Ci[2] - 4. * Ci[6]); D[3] = self[1] * (Ci[3] - 4. * Ci[7]);
            D[4] = self[1] * (Ci[4] - 4. * Ci[4]); D[5] = self[1] * (Ci[5] - 4. * Ci[1]);
            D[6] = self[1] * (Ci[6] - 4. * Ci[2]); D[7] = self[1] * (Ci[7] - 4. * Ci[3]);
            for (int j = 0; j < N; ++j) {
                const double d0[3] = {(double)pos4[4 * i] - pos4[4 * j], (double)pos4[4 * i + 1] - pos4[4 * j + 1],
                                      (double)pos4[4 * i + 2] - pos4[4 * j + 2]};
                for (int n3 = -n3m; n3 <= n3m; ++n3)
                    for (int n2 = -n2m; n2 <= n2m; ++n2)
                        for (int n1 = -n1m; n1 <= n1m; ++n1) {
                            if (i == j && n1 == 0 && n2 == 0 && n3 == 0) continue;
                            const double R0[3] = {d0[0] + n1 * Lx + n2 * xy * Ly, d0[1] + n2 * Ly, d0[2] + n3 * Lz};
                            const double r2 = R0[0] * R0[0] + R0[1] * R0[1] + R0[2] * R0[2];
                            if (r2 >= r_cut * r_cut || r2 == 0.0) continue;
                            const double dist = sqrt(r2);
                            double fv[7];
                            fsd_ewald_functions(dist, xi, fv);
                            /* the table's stored UC entries are g1 / 2 and -g2 / 2, used sign-flipped */
                            const double f1 = fv[0], f2 = fv[1], g1 = -fv[2] / 2, g2 = fv[3] / 2;
                            const double h1 = fv[4], h2 = fv[5], h3 = fv[6];
                            const double R[3] = {R0[0] / dist, R0[1] / dist, R0[2] / dist};
                            const double *Fj = F3 + 3 * j, *Cj = C8 + 8 * j;
                            const double rdotf = dot3(R, Fj);
                            const double Cm[3][3] = {{Cj[0], Cj[1], Cj[2]}, {Cj[5], Cj[4], Cj[3]},
                                                     {Cj[6], Cj[7], -(Cj[0] + Cj[4])}};
                            double Cdr[3], rdC[3];
                            for (int a = 0; a < 3; ++a) {
                                Cdr[a] = Cm[a][0] * R[0] + Cm[a][1] * R[1] + Cm[a][2] * R[2];
                                rdC[a] = R[0] * Cm[0][a] + R[1] * Cm[1][a] + R[2] * Cm[2][a];
                            }
                            const double rCr = dot3(R, Cdr);
                            for (int a = 0; a < 3; ++a) {
                                u[a] += f1 * Fj[a] + (f2 - f1) * rdotf * R[a];
                                u[a] += g1 * (Cdr[a] - rCr * R[a]) + g2 * (rdC[a] - 4. * rCr * R[a]);
                            }
                            static const int A[8] = {0, 0, 0, 1, 1, 1, 2, 2};
                            static const int B[8] = {0, 1, 2, 2, 1, 0, 0, 1};
                            static const int Cs[8] = {0, 5, 6, 7, 4, 1, 2, 3};
                            for (int c = 0; c < 8; ++c) {
                                const int a = A[c], bb = B[c];
                                const double del = (a == bb) ? 1.0 : 0.0;
                                D[c] += -g1 * (R[a] * Fj[bb] - rdotf * R[a] * R[bb])
                                        - g2 * (del * rdotf + Fj[a] * R[bb] - 4. * rdotf * R[a] * R[bb]);
                                D[c] += h1 * (Cj[c] - 4. * Cj[Cs[c]]) + h2 * (R[a] * Cdr[bb] - rCr * R[a] * R[bb])
                                        + h3 * (del * rCr + Cdr[a] * R[bb] + R[a] * rdC[bb] + rdC[a] * R[bb]
                                                - 6. * rCr * R[a] * R[bb] - Cj[Cs[c]]);
                            }
                        }
            }
            for (int a = 0; a < 3; ++a) u3[3 * i + a] += u[a];
            for (int c = 0; c < 8; ++c) D8[8 * i + c] += D[c];
        }
    }
    if (parts & 2) {
        /* reciprocal vectors of the (sheared) box, the construction of gridk_value with the exact 2 pi */
        const int nxm = (int)ceil(k_cut * Lx / (2 * M_PI)) + 1, nym = (int)ceil(k_cut * Ly / (2 * M_PI) + fabs(xy) * nxm * Ly / Lx) + 1;
        const int nzm = (int)ceil(k_cut * Lz / (2 * M_PI)) + 1;
        const int nth = omp_get_max_threads();
        double *acc = calloc((size_t)nth * 11 * N, sizeof(double));
        const double xisq = xi * xi;
#pragma omp parallel
        {
            double *my = acc + (size_t)omp_get_thread_num() * 11 * N;
            double complex *ph = malloc(sizeof(double complex) * N);
#pragma omp for schedule(dynamic, 1) collapse(2)
            for (int nx = -nxm; nx <= nxm; ++nx)
                for (int ny = -nym; ny <= nym; ++ny)
                    for (int nz = -nzm; nz <= nzm; ++nz) {
                        const double kx = 2 * M_PI * nx / Lx, ky = 2 * M_PI * (ny - xy * nx * Ly / Lx) / Ly;
                        const double kz = 2 * M_PI * nz / Lz;
                        const double k2 = kx * kx + ky * ky + kz * kz;
                        if (k2 == 0.0 || k2 > k_cut * k_cut) continue;
                        const double kv[3] = {kx, ky, kz};
                        const double w = 6.0 * M_PI * (1.0 + k2 / 4.0 / xisq) * exp(-k2 / 4.0 / xisq) / k2 / V;
                        /* structure factors F(k) = sum_j F_j e^{-i k.x_j}, C(k) likewise */
                        double complex f[3] = {0, 0, 0}, cc[8] = {0};
                        for (int j = 0; j < N; ++j) {
                            const double kr = kx * pos4[4 * j] + ky * pos4[4 * j + 1] + kz * pos4[4 * j + 2];
                            ph[j] = cos(kr) + I * sin(kr);
                            const double complex e = conj(ph[j]);
                            for (int a = 0; a < 3; ++a) f[a] += F3[3 * j + a] * e;
                            for (int c = 0; c < 8; ++c) cc[c] += C8[8 * j + c] * e;
                        }
                        /* Green's function of Mobility.cu:315-479 on (f, C) */
                        const double complex Cm[3][3] = {{cc[0], cc[1], cc[2]}, {cc[5], cc[4], cc[3]},
                                                         {cc[6], cc[7], -(cc[0] + cc[4])}};
                        const double kk = sqrt(k2);
                        const double complex kdF = (kv[0] * f[0] + kv[1] * f[1] + kv[2] * f[2]) / k2;
                        double complex Cdk[3];
                        for (int a = 0; a < 3; ++a) Cdk[a] = Cm[a][0] * kv[0] + Cm[a][1] * kv[1] + Cm[a][2] * kv[2];
                        const double complex kCk = (Cdk[0] * kv[0] + Cdk[1] * kv[1] + Cdk[2] * kv[2]) / k2;
                        const double sk = sin(kk), ck = cos(kk);
                        const double j0 = sk / kk, j1t = 3.0 * (sk - kk * ck) / (k2 * kk);
                        const double Buf = w * j0 * j0, Buc = -w * j0 * j1t, Bdf = w * j0 * j1t;
                        const double Bdc = w * 9.0 * ((sk - kk * ck) / (k2 * kk)) * ((sk - kk * ck) / (k2 * kk));
                        double complex ou[3], od[3][3];
                        for (int a = 0; a < 3; ++a) ou[a] = (f[a] - kv[a] * kdF) * Buf + I * (Cdk[a] - kv[a] * kCk) * Buc;
                        for (int a = 0; a < 3; ++a)
                            for (int c = 0; c < 3; ++c)
                                od[a][c] = I * (f[a] * kv[c] - kv[a] * kv[c] * kdF) * Bdf
                                           + (Cdk[a] * kv[c] - kv[a] * kv[c] * kCk) * Bdc;
                        /* gradient components in the stored order XX XY XZ YZ YY YX ZX ZY, then the
                         * contraction's transposed slots (Mobility.cu:832-833) as wave_space writes them */
                        const double complex g8[8] = {od[0][0], od[0][1], od[0][2], od[1][2], od[1][1], od[1][0],
                                                      od[2][0], od[2][1]};
                        const int slot[8] = {0, 5, 6, 7, 4, 1, 2, 3};
                        for (int i = 0; i < N; ++i) {
                            const double complex p = ph[i];
                            for (int a = 0; a < 3; ++a) my[11 * (size_t)i + a] += creal(ou[a] * p);
                            for (int c = 0; c < 8; ++c) my[11 * (size_t)i + 3 + slot[c]] += creal(g8[c] * p);
                        }
                    }
            free(ph);
        }
        for (int t = 0; t < nth; ++t)
            for (int i = 0; i < N; ++i) {
                const double *a = acc + (size_t)t * 11 * N + 11 * (size_t)i;
                for (int c = 0; c < 3; ++c) u3[3 * i + c] += a[c];
                for (int c = 0; c < 8; ++c) D8[8 * i + c] += a[3 + c];
            }
        free(acc);
    }
    for (int i = 0; i < N; ++i) {
        double W[3], E5[5];
        d2we(D8 + 8 * i, W, E5);
        for (int a = 0; a < 3; ++a) { gU[6 * i + a] = u3[3 * i + a]; gU[6 * i + 3 + a] = W[a]; }
        for (int c = 0; c < 5; ++c) gU[6 * (size_t)N + 5 * i + c] = E5[c];
    }
    free(F3); free(C8); free(u3); free(D8);
}

/*
 * Real-space Ewald coefficient table of the PSE-split RPY grand mobility
 * (force-torque-stresslet level), for spheres of radius a = 1.
 *
 * Follows Stokes::setParams (Stokes.cc:344-678): entries r_k = (k+1)*dr for
 * k = 0..nR-1, each stored as two float4 = 8 floats
 *     [2k]   = { f1 (I - rr), f2 (rr), g1/2, -g2/2 }   (UF and UC couplings)
 *     [2k+1] = { h1, h2, h3, 0 }                       (DC coupling)
 * The closed forms (Fiore & Swan, J. Fluid Mech. 878, 544 (2019)) contain
 * large cancelling terms at small r, so they are evaluated in binary128 and
 * rounded to float once at the end, exactly as the reference does.  The
 * constants dr and pi are the double-precision literals the reference widens
 * to binary128 (Stokes.cc:359, 383), so the float table is reproduced.
 *
 * r == 2a exactly is never a table node (dr is the binary double nearest
 * 0.001, so (k+1)*dr != 2 in binary128), hence only the r > 2a and r < 2a
 * branches are needed (Stokes.cc:412 and :573).
 *
 * Compiled with gcc and libquadmath into the library and, unchanged, into the
 * oracle (this formula is shared; it is pinned by quadrature tests, not by a
 * second transcription).
 */
#include <quadmath.h>
#include <math.h>

typedef __float128 Q;

/* one table node: out[0..6] = f1, f2, g1, g2, h1, h2, h3 (binary128) */
static void ewald_node(Q R, Q X, Q sp, Q out[7])
{
    const Q R2 = R * R, R3 = R2 * R, R4 = R2 * R2, R5 = R4 * R, R6 = R3 * R3, R8 = R4 * R4;
    const Q X2 = X * X, X3 = X2 * X, X4 = X2 * X2, X5 = X4 * X, X6 = X3 * X3, X7 = X6 * X, X8 = X4 * X4;
    const Q Rp = R + 2, Rm = R - 2;           /* r + 2a, r - 2a */
    const Q Rp2 = Rp * Rp, Rm2 = Rm * Rm;
    const Q e0 = expq(-(R2 * X2));
    const Q ep = expq(-(Rp2 * X2));
    const Q em = expq(-(Rm2 * X2));
    const Q c0 = erfcq(R * X);
    const Q cp = erfcq(Rp * X);
    const Q cm = erfcq(Rm * X);              /* erfc((r - 2a) xi) */
    const Q cn = erfcq(2 * X - R * X);       /* erfc((2a - r) xi) */

    /* ---- pieces common to both branches ---- */
    /* f1 */
    const Q f1_c0 = 3 * c0 * (-12 * R4 + 1 / X4) / (128 * R3);
    const Q f1_e0 = 3 * e0 * sp * (1 + 6 * R2 * X2) / (64 * R2 * X3);
    const Q f1_ep = ep * sp * (8 * R * X2 - 16 * X2 + (2 - 28 * R2 * X2) - 3 * (R + 6 * R3 * X2)) / (128 * R3 * X3);
    const Q f1_em = em * sp * (8 * R * X2 + 16 * X2 + (-2 + 28 * R2 * X2) - 3 * (R + 6 * R3 * X2)) / (128 * R3 * X3);
    const Q f1_pp = 128 + 64 / R3 + 96 / R + 36 * R - 3 / (R3 * X4);   /* multiplies cp/256 in both branches */
    /* f2 */
    const Q f2_e0 = 3 * e0 * sp * (-1 + 2 * R2 * X2) / (32 * R2 * X3);
    const Q f2_em = -(2 + 3 * R) * em * sp * (-1 - 8 * R * X2 + 8 * X2 + 2 * R2 * X2) / (64 * R3 * X3);
    const Q f2_ep = (2 - 3 * R) * ep * sp * (-1 + 8 * R * X2 + 8 * X2 + 2 * R2 * X2) / (64 * R3 * X3);
    const Q f2_c0 = -3 * c0 * (1 + 4 * R4 * X4) / (64 * R3 * X4);
    const Q f2_pp = 64 - 64 / R3 + 96 / R + 12 * R + 3 / (R3 * X4);    /* multiplies cp/128 in both branches */
    /* g1 */
    const Q g1_e0 = e0 * sp * (9 + 15 * R2 * X2 - 30 * R4 * X4) / (64 * R3 * X5);
    const Q g1_ep = ep * sp * (18 - 45 * R - 3 * Rp * (-16 * R + 8 + 25 * R2) * X2
                               + 6 * Rp * (-32 * R + 32 + 44 * R2 - 36 * R3 + 25 * R4) * X4) / (640 * R4 * X5);
    const Q g1_em = em * sp * (-9 * (2 + 5 * R) + 3 * (2 - R) * (16 * R + 8 + 25 * R2) * X2
                               - 6 * (2 - R) * (32 * R + 32 + 44 * R2 + 36 * R3 + 25 * R4) * X4) / (640 * R4 * X5);
    const Q g1_c0 = 3 * c0 * (3 + 3 * R2 * X2 + 20 * R6 * X6) / (128 * R4 * X6);
    const Q g1_pm = (15 + 5 * R2 * X2 * (3 + 64 * X4) + 512 * X6 - 256 * R5 * X6 + 100 * R6 * X6) / (1280 * R4 * X6);
    const Q g1_cp = -3 * cp * (15 + 5 * R2 * X2 * (3 + 64 * X4) + 512 * X6 + 256 * R5 * X6 + 100 * R6 * X6) / (1280 * R4 * X6);
    /* g2 */
    const Q g2_e0 = -3 * e0 * sp * (3 - R2 * X2 + 2 * R4 * X4) / (64 * R3 * X5);
    const Q g2_pm = 24 * R + 16 + 14 * R2 + 5 * R3;
    const Q g2_em = em * sp * (18 + 45 * R - 3 * g2_pm * X2 + 6 * g2_pm * Rm2 * X4) / (640 * R4 * X5);
    const Q g2_pp = -24 * R + 16 + 14 * R2 - 5 * R3;
    const Q g2_ep = ep * sp * (-18 + 45 * R + 3 * g2_pp * X2 - 6 * g2_pp * Rp2 * X4) / (640 * R4 * X5);
    const Q g2_cm = 3 * cm * (15 - 15 * R2 * X2 + 4 * (128 - 80 * R2 + 16 * R5 - 5 * R6) * X6) / (1280 * R4 * X6);
    const Q g2_c0 = 3 * c0 * (-3 + 3 * R2 * X2 + 4 * R6 * X6) / (128 * R4 * X6);
    const Q g2_cp = -3 * cp * (-15 + 15 * R2 * X2 + 4 * (-128 + 80 * R2 + 16 * R5 + 5 * R6) * X6) / (1280 * R4 * X6);
    /* h1 */
    const Q h1_e0 = 3 * e0 * sp * (27 - 2 * X2 * (15 * R2 + 2 * R4 * X2 - 4 * R6 * X4
                                  + 48 * (3 - R2 * X2 + 2 * R4 * X4))) / (4096 * R4 * X7);
    const Q h1_qm = 96 * R + 48 + 80 * R2 + 40 * R3 + 5 * R4;
    const Q h1_qp = -96 * R + 48 + 80 * R2 - 40 * R3 + 5 * R4;
    const Q h1_em = 3 * em * sp * (270 - 135 * R + 6 * (2 + 5 * R) * (12 + 5 * R2) * X2
                                   - 4 * (144 * R + 96 + 64 * R2 - 30 * R4 - 5 * R5) * X4
                                   + 8 * (2 - R) * (2 - R) * (2 - R) * h1_qm * X6) / (40960 * R5 * X7);
    const Q h1_ep = 3 * ep * sp * (-135 * Rp - 6 * (2 - 5 * R) * (12 + 5 * R2) * X2
                                   + 4 * (-144 * R + 96 + 64 * R2 - 30 * R4 + 5 * R5) * X4
                                   - 8 * h1_qp * Rp * Rp2 * X6) / (40960 * R5 * X7);
    const Q h1_c0 = 3 * c0 * (27 + 8 * X2 * (-6 * R2 + 9 * R4 * X2 - 2 * R8 * X6
                                             + 12 * (-3 + 3 * R2 * X2 + 4 * R6 * X6))) / (8192 * R5 * X8);
    const Q h1_base = -135 + 240 * (6 + R2) * X2 - 360 * R2 * (4 + R2) * X4;
    const Q h1_cm = 3 * cm * (h1_base + 16 * h1_qm * Rm2 * Rm2 * X8) / (81920 * R5 * X8);
    const Q h1_cp = 3 * cp * (h1_base + 16 * h1_qp * Rp2 * Rp2 * X8) / (81920 * R5 * X8);
    /* h2 */
    const Q h2_e0 = 9 * e0 * sp * (-45 - 78 * R2 * X2 + 28 * R4 * X4 + 32 * X2 * (15 + 19 * R2 * X2 + 10 * R4 * X4)
                                   - 56 * R6 * X6) / (4096 * R4 * X7);
    const Q h2_ep = 9 * ep * sp * (45 * Rp + 6 * (-20 * R + 8 + 46 * R2 + 13 * R3) * X2
                                   - 4 * Rp * (-32 * R + 16 + 48 * R2 - 56 * R3 + 7 * R4) * X4
                                   + 8 * Rp * (16 + 16 * R2 + 7 * R4) * Rm2 * X6) / (8192 * R5 * X7);
    const Q h2_em = 9 * em * sp * (45 * Rm - 6 * (20 * R + 8 + 46 * R2 - 13 * R3) * X2
                                   + 4 * (2 - R) * (32 * R + 16 + 48 * R2 + 56 * R3 + 7 * R4) * X4
                                   - 8 * (2 - R) * (16 + 16 * R2 + 7 * R4) * Rp2 * X6) / (8192 * R5 * X7);
    const Q h2_poly = -45 + 8 * X2 * (60 - 6 * R2 + 9 * R2 * (4 + R2) * X2 + 2 * (256 + 128 * R2 - 40 * R6 + 7 * R8) * X6);
    const Q h2_cp = -9 * cp * h2_poly / (16384 * R5 * X8);
    const Q h2_c0 = -9 * c0 * (45 + 8 * X2 * (6 * R2 - 9 * R4 * X2 - 14 * R8 * X6
                                              + 4 * (-15 - 9 * R2 * X2 + 20 * R6 * X6))) / (8192 * R5 * X8);
    /* h3 */
    const Q h3_e0 = 9 * e0 * sp * (-45 + 18 * R2 * X2 - 4 * R4 * X4 + 32 * X2 * (15 + R2 * X2 - 2 * R4 * X4)
                                   + 8 * R6 * X6) / (4096 * R4 * X7);
    const Q tm3 = (2 - R) * (2 - R) * (2 - R);
    const Q h3_ep = 9 * ep * sp * (45 * Rp + 6 * (2 - 3 * R) * Rm2 * X2 - 4 * tm3 * (4 + R2) * X4
                                   + 8 * tm3 * (4 + R2) * Rp2 * X6) / (8192 * R5 * X7);
    const Q h3_em = 9 * em * sp * (45 * Rm - 6 * (2 + 3 * R) * Rp2 * X2 + 4 * (4 + R2) * Rp * Rp2 * X4
                                   - 8 * (4 + R2) * Rm2 * Rp * Rp2 * X6) / (8192 * R5 * X7);
    const Q d4 = 4 - R2;
    const Q h3_poly = -45 + 8 * X2 * (60 + 6 * R2 - 3 * R2 * (12 + R2) * X2 + 2 * (4 + R2) * d4 * d4 * d4 * X6);
    const Q h3_cmp = -9 * (cm + cp) * h3_poly / (16384 * R5 * X8);
    const Q h3_c0 = 9 * c0 * (-45 + 8 * X2 * (6 * R2 - 3 * R4 * X2 - 2 * R8 * X6
                                              + 4 * (15 - 9 * R2 * X2 + 4 * R6 * X6))) / (8192 * R5 * X8);

    if (R > 2) {
        out[0] = -1 + 1 / (2 * R3) + 3 / (4 * R) + f1_c0 + (9 * R / 32 - 3 / (128 * R3 * X4))
               + cp * f1_pp / 256
               + cn * (128 - 64 / R3 - 96 / R - 36 * R + 3 / (R3 * X4)) / 256
               + f1_e0 + f1_ep + f1_em;
        out[1] = -1 - 1 / R3 + 3 / (2 * R) + 3 * (4 * R4 + 1 / X4) / (64 * R3)
               + cn * (64 + 64 / R3 - 96 / R - 12 * R - 3 / (R3 * X4)) / 128
               + cp * f2_pp / 128
               + f2_e0 + f2_em + f2_ep + f2_c0;
        out[2] = g1_e0 + g1_ep + g1_em + g1_c0 - 3 * cm * g1_pm + g1_cp;
        out[3] = g2_e0 + g2_em + g2_ep + g2_cm + g2_c0 + g2_cp;
        out[4] = h1_e0 + h1_em + h1_ep + h1_c0 + h1_cm + h1_cp;
        out[5] = h2_e0 + h2_ep + h2_em - 9 * cm * h2_poly / (16384 * R5 * X8) + h2_cp + h2_c0;
        out[6] = h3_e0 + h3_ep + h3_em + h3_cmp + h3_c0;
    } else {
        /* overlapping pair: piecewise-polynomial RPY overlap terms added */
        out[0] = -9 * R / 32 + 1 - 1 / (2 * R3) - 3 / (4 * R) + f1_c0
               + cm * (-128 + 64 / R3 + 96 / R + 36 * R - 3 / (R3 * X4)) / 256
               + cp * f1_pp / 256
               + f1_e0 + f1_ep + f1_em;
        out[1] = (2 + 3 * R) * tm3 / (16 * R3)
               + cm * (-64 - 64 / R3 + 96 / R + 12 * R + 3 / (R3 * X4)) / 128
               + cp * f2_pp / 128
               + f2_e0 + f2_em + f2_ep + f2_c0;
        out[2] = -9 / (128 * R4 * X6) - 9 / (128 * R2 * X4)
               + g1_e0 + g1_ep + g1_em + g1_c0 + 3 * cn * g1_pm + g1_cp;
        out[3] = -3 * R / 10 - 12 / (5 * R4) + 3 / (2 * R2) + 3 * R2 / 32
               + g2_e0 + g2_em + g2_ep + g2_cm + g2_c0 + g2_cp;
        out[4] = 9 * R / 64 - (Q)3 / 10 - 9 / (10 * R5) + 3 / (4 * R3) - 3 * R3 / 512
               + h1_e0 + h1_em + h1_ep + h1_c0 + h1_cm + h1_cp;
        out[5] = 63 * R / 64 - (Q)3 / 2 + 9 / (2 * R5) - 3 / (4 * R3) - 33 * R3 / 512
               + 9 / (128 * R3 * X6) - 27 / (64 * R3 * X4)
               + h2_e0
               + 3 * cn * (-3 + 18 * X2 * (1 - 4 * R4 * X4) + 128 * X6 + 64 * R3 * X6 + 8 * R6 * X6) / (256 * R3 * X6)
               + h2_ep + h2_em + h2_cp
               + 3 * cm * (135 + 8 * X2 * (-6 * (30 + R2) + 9 * (4 - 3 * R2) * R2 * X2
                                           + 2 * (-768 + 128 * R2 + 256 * R5 - 168 * R6 + 11 * R8) * X6)) / (16384 * R5 * X8)
               + h2_c0;
        out[6] = 9 * R / 64 + 9 / (2 * R5) - 9 / (4 * R3) - 9 * R3 / 512
               + h3_e0 + h3_ep + h3_em + h3_cmp + h3_c0;
    }
}

/* Number of table nodes for a given cutoff, as Stokes.cc:350-351, 372. */
int fsd_ewald_table_size(float ewald_cut, float ewald_dr)
{
    int ewald_n = (int)(ewald_cut / ewald_dr - 1.0f);
    return ewald_n + 1;
}

/* Fill nR nodes (8 floats each) for splitting parameter xi. */
void fsd_ewald_table_fill(float xi, int nR, float *tab)
{
    const Q dr = (Q)(double)0.001;
    const Q pi_d = (Q)(double)3.1415926535897932384626433832795;
    const Q sp = powq(pi_d, (Q)-0.5);
    const Q X = (Q)xi;
    for (int k = 0; k < nR; ++k) {
        Q v[7];
        const Q R = (Q)k * dr + dr;
        ewald_node(R, X, sp, v);
        float *e = tab + 8 * k;
        e[0] = (float)v[0];
        e[1] = (float)v[1];
        e[2] = (float)(v[2] / 2);
        e[3] = (float)(-v[3] / 2);
        e[4] = (float)v[4];
        e[5] = (float)v[5];
        e[6] = (float)v[6];
        e[7] = 0.0f;
    }
}

/* Same functions in double at an arbitrary distance (used by tests and the
 * oracle's exact-table variant): out = f1, f2, g1, g2, h1, h2, h3. */
void fsd_ewald_functions(double r, double xi, double out[7])
{
    const Q pi_d = (Q)(double)3.1415926535897932384626433832795;
    Q v[7];
    ewald_node((Q)r, (Q)xi, powq(pi_d, (Q)-0.5), v);
    for (int i = 0; i < 7; ++i) out[i] = (double)v[i];
}

/* Ewald self terms (Stokes.cc:362-371), evaluated with the reference's mixed
 * float/double expression so the float results agree. */
void fsd_ewald_self(float xi, float self_out[2])
{
    const float pi12 = 1.77245385091f;
    const float pi = 3.1415926536f;
    const float aa = 1.0f;
    const float axi = aa * xi;
    const float axi2 = axi * axi;
    self_out[0] = (float)((1. + 4. * pi12 * axi * erfcf((float)(2. * axi)) - expf((float)(-4. * axi2)))
                          / (4. * pi12 * axi * aa));
    self_out[1] = (float)((-3. * erfcf((float)(2. * aa * xi)) * powf(aa, -3.f)) / 10.
                          - (3. * powf(aa, -6.f) * powf(pi, -0.5f) * powf(xi, -3.f)) / 80.
                          - (9. * powf(aa, -4.f) * powf(pi, -0.5f) * powf(xi, -1.f)) / 40.
                          + (3. * expf((float)(-4. * powf(aa, 2.f) * powf(xi, 2.f))) * powf(aa, -6.f)
                             * powf(pi, -0.5f) * powf(xi, -3.f) * (1. + 10. * powf(aa, 2.f) * powf(xi, 2.f))) / 80.);
}

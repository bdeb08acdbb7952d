// Host FP64 part of the Lanczos square roots of the Brownian sampler
// (Sqrt_multiply, Helper_Brownian.cu:476-583): t = T^(1/2) e_1 for the m x m
// symmetric tridiagonal Lanczos matrix T, through its eigendecomposition
// T = Q diag(lambda) Q^T computed with cyclic Jacobi rotations (Golub and
// Van Loan, Matrix Computations, 4th ed., section 8.5.2) in double precision.
#include "lanczos_host.h"

#include <cmath>
#include <vector>

// A (m x m, row-major, symmetric): eigenvalues end on the diagonal, the
// eigenvectors in the columns of Q
static void jacobi_eigen(int m, std::vector<double> &A, std::vector<double> &Q)
{
    Q.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) Q[(size_t)i * m + i] = 1.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int p = 0; p < m; ++p)
            for (int q = 0; q < m; ++q) {
                const double v = A[(size_t)p * m + q] * A[(size_t)p * m + q];
                tot += v;
                if (p != q) off += v;
            }
        if (off <= 1e-32 * tot) break;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                const double apq = A[(size_t)p * m + q];
                if (std::fabs(apq) <= 1e-300) continue;
                const double tau = (A[(size_t)q * m + q] - A[(size_t)p * m + p]) / (2.0 * apq);
                const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < m; ++k) {      // A <- A J
                    const double akp = A[(size_t)k * m + p], akq = A[(size_t)k * m + q];
                    A[(size_t)k * m + p] = c * akp - s * akq;
                    A[(size_t)k * m + q] = s * akp + c * akq;
                }
                for (int k = 0; k < m; ++k) {      // A <- J^T A
                    const double apk = A[(size_t)p * m + k], aqk = A[(size_t)q * m + k];
                    A[(size_t)p * m + k] = c * apk - s * aqk;
                    A[(size_t)q * m + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < m; ++k) {      // Q <- Q J
                    const double qkp = Q[(size_t)k * m + p], qkq = Q[(size_t)k * m + q];
                    Q[(size_t)k * m + p] = c * qkp - s * qkq;
                    Q[(size_t)k * m + q] = s * qkp + c * qkq;
                }
            }
    }
}

double tridiag_sqrt_e1(int m, const double *alpha, const double *off, double *t)
{
    std::vector<double> A((size_t)m * m, 0.0), Q;
    for (int i = 0; i < m; ++i) A[(size_t)i * m + i] = alpha[i];
    for (int i = 0; i + 1 < m; ++i) A[(size_t)i * m + i + 1] = A[(size_t)(i + 1) * m + i] = off[i];
    jacobi_eigen(m, A, Q);
    double lmin = 1e300;
    std::vector<double> s(m);
    for (int j = 0; j < m; ++j) {
        const double lam = A[(size_t)j * m + j];
        lmin = std::fmin(lmin, lam);
        s[j] = std::sqrt(std::fmax(lam, 0.0)) * Q[j];       // lambda_j^(1/2) (Q^T e_1)_j
    }
    for (int i = 0; i < m; ++i) {
        double acc = 0.0;
        for (int j = 0; j < m; ++j) acc += Q[(size_t)i * m + j] * s[j];
        t[i] = acc;
    }
    return lmin;
}

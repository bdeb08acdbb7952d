// Checks tridiag_sqrt_e1 (fsd_amd/csrc/lanczos_host.cpp).  With
// t = T^(1/2) e_1 for symmetric positive definite tridiagonal T:
//   t^T t   = (T)_11      and      t^T T t = (T^2)_11,
// and for a diagonal T, t = sqrt(T_11) e_1.  A wrong matrix function (T e_1,
// T^(-1/2) e_1, a transposed eigenvector matrix) fails at least one of them.
#include "../../fsd_amd/csrc/lanczos_host.h"
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

int main()
{
    std::mt19937 rng(7);
    std::uniform_real_distribution<double> U(0.1, 2.0), V(-0.5, 0.5);
    double worst = 0.0;
    for (int m : {1, 2, 3, 10, 40, 99}) {
        std::vector<double> a(m), o(m > 1 ? m - 1 : 1), t(m), Tt(m, 0.0);
        for (int i = 0; i < m; ++i) a[i] = 2.0 + U(rng);
        for (int i = 0; i + 1 < m; ++i) o[i] = V(rng);
        tridiag_sqrt_e1(m, a.data(), o.data(), t.data());
        double tt = 0.0, tTt = 0.0;
        for (int i = 0; i < m; ++i) {
            Tt[i] = a[i] * t[i] + (i > 0 ? o[i - 1] * t[i - 1] : 0.0) + (i + 1 < m ? o[i] * t[i + 1] : 0.0);
            tt += t[i] * t[i];
        }
        for (int i = 0; i < m; ++i) tTt += t[i] * Tt[i];
        const double T2_11 = a[0] * a[0] + (m > 1 ? o[0] * o[0] : 0.0);
        worst = std::fmax(worst, std::fabs(tt - a[0]) / a[0]);
        worst = std::fmax(worst, std::fabs(tTt - T2_11) / T2_11);
    }
    double d[3] = {4.0, 9.0, 1.0}, z[2] = {0.0, 0.0}, t[3];
    tridiag_sqrt_e1(3, d, z, t);
    worst = std::fmax(worst, std::fabs(t[0] - 2.0) + std::fabs(t[1]) + std::fabs(t[2]));
    printf("max rel error %.3e\n", worst);
    return worst < 1e-12 ? 0 : 1;
}

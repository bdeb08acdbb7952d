// Host FP64 helpers of the Lanczos square roots.
#pragma once

// t = T^(1/2) e_1 for the symmetric tridiagonal T with diagonal alpha[0..m)
// and off-diagonal off[0..m-1) (T_{i,i+1} = off[i]); negative eigenvalues
// (rounding) are clamped to zero.  Returns the smallest eigenvalue of T.
double tridiag_sqrt_e1(int m, const double *alpha, const double *off, double *t);

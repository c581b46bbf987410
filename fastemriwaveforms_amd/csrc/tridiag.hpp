// Elimination of a tridiagonal matrix as LAPACK dgtsv performs it (Gaussian
// elimination with partial pivoting), recorded so that any number of
// right-hand sides can be swept afterwards with exactly dgtsv's arithmetic.
//
// dgtsv's pivot choices and multipliers depend on the matrix only; with the
// recorded step i (pivot flag, multiplier) a right-hand side b is swept as
//   no pivot: b[i+1] = b[i+1] - fact * b[i]
//   pivot:    (b[i], b[i+1]) = (b[i+1], b[i] - fact * b[i+1])
// and back-substituted with the eliminated d, du and the fill-in dl -- the
// same operations dgtsv applies when it carries b along.
#pragma once

#include <math.h>

namespace few {

// In: d (n), du (n-1: A[i][i+1]), dl (n-1: A[i+1][i]).  Out: d, du, dl
// eliminated (dl becomes the second super-diagonal fill-in), piv[i] in
// {0, 1} and fact[i] for i = 0..n-2.  Returns dgtsv's INFO.
__host__ __device__ inline int dgtsv_factor(int n, double *d, double *du, double *dl, double *piv,
                                            double *fact)
{
#pragma clang fp contract(off)
    double f, temp;
    int i;
    for (i = 0; i < n - 2; i++) {
        if (fabs(d[i]) >= fabs(dl[i])) {
            if (d[i] == 0.0) return i + 1;
            f = dl[i] / d[i];
            d[i + 1] = d[i + 1] - f * du[i];
            dl[i] = 0.0;
            piv[i] = 0.0;
        } else {
            f = d[i] / dl[i];
            d[i] = dl[i];
            temp = d[i + 1];
            d[i + 1] = du[i] - f * temp;
            dl[i] = du[i + 1];
            du[i + 1] = -f * dl[i];
            du[i] = temp;
            piv[i] = 1.0;
        }
        fact[i] = f;
    }
    if (n > 1) {
        i = n - 2;
        if (fabs(d[i]) >= fabs(dl[i])) {
            if (d[i] == 0.0) return i + 1;
            f = dl[i] / d[i];
            d[i + 1] = d[i + 1] - f * du[i];
            piv[i] = 0.0;
        } else {
            f = d[i] / dl[i];
            d[i] = dl[i];
            temp = d[i + 1];
            d[i + 1] = du[i] - f * temp;
            du[i] = temp;
            piv[i] = 1.0;
        }
        fact[i] = f;
    }
    if (d[n - 1] == 0.0) return n;
    return 0;
}

}  // namespace few

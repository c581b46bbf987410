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
    // di / dui: the current (already updated) d[i] and du[i], carried in
    // registers so each step waits only on its own arithmetic
    double di = d[0], dui = n > 1 ? du[0] : 0.0;
    double f, temp;
    int i;
    for (i = 0; i < n - 2; i++) {
        const double dli = dl[i], dn = d[i + 1], dun = du[i + 1];
        if (fabs(di) >= fabs(dli)) {
            if (di == 0.0) {
                d[i] = di;
                du[i] = dui;
                return i + 1;
            }
            f = dli / di;
            d[i] = di;
            di = dn - f * dui;
            dl[i] = 0.0;
            du[i] = dui;
            dui = dun;
            piv[i] = 0.0;
        } else {
            f = di / dli;
            d[i] = dli;
            temp = dn;
            di = dui - f * temp;
            dl[i] = dun;
            dui = -f * dun;
            du[i] = temp;
            piv[i] = 1.0;
        }
        fact[i] = f;
    }
    if (n > 1) {
        i = n - 2;
        const double dli = dl[i], dn = d[i + 1];
        if (fabs(di) >= fabs(dli)) {
            if (di == 0.0) {
                d[i] = di;
                du[i] = dui;
                return i + 1;
            }
            f = dli / di;
            d[i] = di;
            di = dn - f * dui;
            du[i] = dui;
            piv[i] = 0.0;
        } else {
            f = di / dli;
            d[i] = dli;
            temp = dn;
            di = dui - f * temp;
            du[i] = temp;
            piv[i] = 1.0;
        }
        fact[i] = f;
    }
    d[n - 1] = di;
    if (di == 0.0) return n;
    return 0;
}

// dgtsv_factor with the operands of the next PF steps (dl[i], d[i+1],
// du[i+1]; no earlier step writes them) read ahead of the dependent chain,
// and each step's two pivot branches as selects around one division: for a
// matrix in LDS, one wait per PF steps and no branch per step on the chain.
// Same operations on the same operands, same results and early returns.
template <int PF>
__host__ __device__ inline int dgtsv_factor_ahead(int n, double *d, double *du, double *dl,
                                                  double *piv, double *fact)
{
#pragma clang fp contract(off)
    double di = d[0], dui = n > 1 ? du[0] : 0.0;
    for (int i0 = 0; i0 < n - 2; i0 += PF) {
        double dlv[PF], dnv[PF], dunv[PF];
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int i = i0 + u < n - 3 ? i0 + u : n - 3;
            dlv[u] = dl[i];
            dnv[u] = d[i + 1];
            dunv[u] = du[i + 1];
        }
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int i = i0 + u;
            if (i < n - 2) {
                const double dli = dlv[u], dn = dnv[u], dun = dunv[u];
                const bool keep = fabs(di) >= fabs(dli);  // no row interchange
                if (keep && di == 0.0) {
                    d[i] = di;
                    du[i] = dui;
                    return i + 1;
                }
                // keep: f = dli / di, d' = dn - f dui, dl = 0, du = dui, dui' = dun
                // swap: f = di / dli, d' = dui - f dn, dl = dun, du = dn, dui' = -f dun
                const double f = (keep ? dli : di) / (keep ? di : dli);
                d[i] = keep ? di : dli;
                dl[i] = keep ? 0.0 : dun;
                du[i] = keep ? dui : dn;
                piv[i] = keep ? 0.0 : 1.0;
                fact[i] = f;
                const double dnew = keep ? dn - f * dui : dui - f * dn;
                dui = keep ? dun : -f * dun;
                di = dnew;
            }
        }
    }
    if (n > 1) {
        const int i = n - 2;
        const double dli = dl[i], dn = d[i + 1];
        double f;
        if (fabs(di) >= fabs(dli)) {
            if (di == 0.0) {
                d[i] = di;
                du[i] = dui;
                return i + 1;
            }
            f = dli / di;
            d[i] = di;
            di = dn - f * dui;
            du[i] = dui;
            piv[i] = 0.0;
        } else {
            f = di / dli;
            d[i] = dli;
            const double temp = dn;
            di = dui - f * temp;
            du[i] = temp;
            piv[i] = 1.0;
        }
        fact[i] = f;
    }
    d[n - 1] = di;
    if (di == 0.0) return n;
    return 0;
}

}  // namespace few

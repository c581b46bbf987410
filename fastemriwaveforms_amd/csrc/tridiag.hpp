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

#define DGTSV_SWEEP_PF 8  // values fetched ahead of the sweep recurrences

// dgtsv's forward sweep of one right-hand side b (stride bs) with the
// elimination F = piv | fact | d | du | dl recorded by dgtsv_factor, then the
// back substitution: the solution into b.  dgtsv's operations in its order;
// the operands of the next DGTSV_SWEEP_PF steps are read before the dependent
// arithmetic and each pivot choice is a select (no branch on the chain).
__host__ __device__ inline void dgtsv_sweeps(int n, const double *F, double *bv, long bs)
{
#pragma clang fp contract(off)
#define SB(i) bv[(long)(i) * bs]
    const double *piv = F, *fact = F + n, *d = F + 2 * n, *du = F + 3 * n, *dl = F + 4 * n;
    double cur = SB(0);
    for (int i0 = 0; i0 < n - 1; i0 += DGTSV_SWEEP_PF) {
        double nbv[DGTSV_SWEEP_PF], fv[DGTSV_SWEEP_PF], pv[DGTSV_SWEEP_PF];
#pragma unroll
        for (int u = 0; u < DGTSV_SWEEP_PF; u++) {
            const int i = i0 + u < n - 2 ? i0 + u : n - 2;
            nbv[u] = SB(i + 1);
            fv[u] = fact[i];
            pv[u] = piv[i];
        }
#pragma unroll
        for (int u = 0; u < DGTSV_SWEEP_PF; u++) {
            const int i = i0 + u;
            if (i < n - 1) {
                // no interchange: b[i] stays, b[i+1] -= f b[i];
                // interchange:    (b[i], b[i+1]) = (b[i+1], b[i] - f b[i+1])
                const double nb = nbv[u];
                const bool keep = pv[u] == 0.0;
                SB(i) = keep ? cur : nb;
                cur = keep ? nb - fv[u] * cur : cur - fv[u] * nb;
            }
        }
    }
    double x_next = cur / d[n - 1];
    SB(n - 1) = x_next;
    // row n - 2 has no second super-diagonal term
    double x_next2 = x_next;
    x_next = (SB(n - 2) - du[n - 2] * x_next) / d[n - 2];
    SB(n - 2) = x_next;
    for (int i0 = n - 3; i0 >= 0; i0 -= DGTSV_SWEEP_PF) {
        double bvv[DGTSV_SWEEP_PF], duv[DGTSV_SWEEP_PF], dlv[DGTSV_SWEEP_PF], dv[DGTSV_SWEEP_PF];
#pragma unroll
        for (int u = 0; u < DGTSV_SWEEP_PF; u++) {
            const int ii = i0 - u > 0 ? i0 - u : 0;
            bvv[u] = SB(ii);
            duv[u] = du[ii];
            dlv[u] = dl[ii];
            dv[u] = d[ii];
        }
#pragma unroll
        for (int u = 0; u < DGTSV_SWEEP_PF; u++) {
            const int ii = i0 - u;
            if (ii >= 0) {
                const double xi = (bvv[u] - duv[u] * x_next - dlv[u] * x_next2) / dv[u];
                SB(ii) = xi;
                x_next2 = x_next;
                x_next = xi;
            }
        }
    }
#undef SB
}

}  // namespace few

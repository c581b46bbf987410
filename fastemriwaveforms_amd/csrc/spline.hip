// Batched not-a-knot cubic-spline build for gfx950.
//
// Replaces the reference's four-launch sequence fill_B -> cuSPARSE
// gtsv2StridedBatch -> set_spline_constants -> fill_final_derivs
// (src/few/cutils/interpolate.cu:284-325) with ONE fused kernel: one lane
// owns one system, assembles its rows on the fly (interpolate.cu:49-102),
// eliminates them with LAPACK dgtsv's partial-pivoting recurrence (the
// routine the reference CPU backend calls, interpolate.cu:231-236), and
// emits the cubic coefficients (interpolate.cu:28-43, 150-198) plus the
// last-knot re-expansion (interpolate.cu:242-280) during back substitution.
//
// Floating point: this translation unit is compiled with contraction off so
// every operation rounds exactly as the CPU oracle's -ffp-contract=off build;
// phase splines built here therefore match the oracle bit for bit, which the
// FP64 mode sum needs for rel-L2 1e-12 at phases of 1e6-1e7 rad.
//
// Scratch (the reference's B / upper_diag / diag / lower_diag buffers) is
// used knot-major, [knot][system], so the per-lane sequential sweeps are
// coalesced across the wavefront.

#pragma clang fp contract(off)

#include "batch.hpp"
#include "few_common.hpp"
#include "tridiag.hpp"

namespace few {

#define SPLINE_MAX_GROUPS 4
#define SPLINE_PF 8  // prefetch distance (steps) of the sequential sweeps
#define SPLINE_SHARED_MAX_L 150  // LDS path: (2 * 32 + 6) * L doubles per 32-system block
#define SPLINE_TILE_BATCH 24     // knot-value loads in flight per lane while staging a tile

struct SplineArgs {
    SplineGroup g[SPLINE_MAX_GROUPS];
    int ngroups;
    int blk_prefix[SPLINE_MAX_GROUPS + 1];   // first block of each group
    int col_prefix[SPLINE_MAX_GROUPS + 1];   // first scratch column of each group
    int length;
    int nsys_total;
    double *sD, *sDU, *sDL, *sB;
    int *info;
    int shared_lds;      // groups with shared knots take the LDS path (spline_block_shared)
};

// Row i of the not-a-knot system (interpolate.cu:49-102).  xa/ya are the
// three knots the row reads: (x[i-1], x[i], x[i+1]) for interior rows,
// (x0, x1, x2) for row 0 and (x[n-3], x[n-2], x[n-1]) for row n-1.
__host__ __device__ inline void prep_row(int i, int n, double x0, double x1, double x2, double y0,
                                         double y1, double y2, double &b, double &ud, double &dg,
                                         double &ld)
{
    if (i == n - 1) {
        double dx1 = x1 - x0;
        double dx2 = x2 - x1;
        double d = x2 - x0;
        double slope1 = (y1 - y0) / dx1;
        double slope2 = (y2 - y1) / dx2;
        b = ((dx2 * dx2 * slope1 + (2 * d + dx2) * dx1 * slope2) / d);
        dg = dx1;
        ld = d;
        ud = 0.0;
    } else if (i == 0) {
        double dx1 = x1 - x0;
        double dx2 = x2 - x1;
        double d = x2 - x0;
        double slope1 = (y1 - y0) / dx1;
        double slope2 = (y2 - y1) / dx2;
        b = ((dx1 + 2 * d) * dx2 * slope1 + dx1 * dx1 * slope2) / d;
        dg = dx2;
        ud = d;
        ld = 0.0;
    } else {
        double dx1 = x1 - x0;
        double dx2 = x2 - x1;
        double slope1 = (y1 - y0) / dx1;
        double slope2 = (y2 - y1) / dx2;
        b = 3.0 * (dx2 * slope1 + dx1 * slope2);
        dg = 2 * (dx1 + dx2);
        ud = dx1;
        ld = dx2;
    }
}

// The whole computation of system s of group G: n knots, scratch column col
// of ns columns.  Host-callable so the CPU test harness runs this exact code
// against the oracle.
__host__ __device__ void spline_system_core(const SplineGroup &G, int s, int n, long col, long ns,
                                            const SplineScratch &a)
{

    const double *yv = G.y + (long)s * G.y_sys;  // may alias out (in-place build)
    const double *tv = G.t + (long)s * G.t_sys;
    const long yk = G.y_knot, tk = G.t_knot;
#define XK(i) tv[(long)(i) * tk]
#define YK(i) yv[(long)(i) * yk]

    // ---------------- forward elimination (dgtsv, NRHS = 1) ----------------
    // The knot loads are issued SPLINE_PF steps ahead through a register ring
    // so the sequential recurrence does not wait on memory each step; the
    // arithmetic is unchanged.
    double xa = XK(0), xb = XK(1), xc = XK(2);
    double ya = YK(0), yb = YK(1), yc = YK(2);
    double cb, cdu, cd, cld;
    prep_row(0, n, xa, xb, xc, ya, yb, yc, cb, cdu, cd, cld);
    int info = 0;
    int stop = n - 1;  // rows [0, stop) are stored in scratch

    // One dgtsv elimination step i, with row r = i+1 assembled from the
    // current window.  Returns false when LAPACK would stop (zero pivot).
    auto elim = [&](int i) -> bool {
        const int r = i + 1;
        double nb, ndu, nd, nld;
        prep_row(r, n, xa, xb, xc, ya, yb, yc, nb, ndu, nd, nld);
        double Di, DUi, DLi, Bi;
        if (fabs(cd) >= fabs(nld)) {
            if (cd != 0.0) {
                const double fact = nld / cd;
                nd = nd - fact * cdu;
                nb = nb - fact * cb;
            } else {
                info = i + 1;
                a.sB[(long)i * ns + col] = cb;
                a.sB[(long)r * ns + col] = nb;
                stop = r + 1;
                return false;
            }
            Di = cd; DUi = cdu; DLi = (i < n - 2) ? 0.0 : nld; Bi = cb;
        } else {
            const double fact = cd / nld;
            Di = nld;
            const double temp = nd;
            nd = cdu - fact * temp;
            if (i < n - 2) {
                DLi = ndu;
                ndu = -fact * DLi;
            } else {
                DLi = 0.0;
            }
            DUi = temp;
            Bi = nb;
            nb = cb - fact * nb;
        }
        a.sD[(long)i * ns + col] = Di;
        a.sDU[(long)i * ns + col] = DUi;
        a.sDL[(long)i * ns + col] = DLi;
        a.sB[(long)i * ns + col] = Bi;
        cd = nd; cdu = ndu; cb = nb;
        return true;
    };

    // row 1 reads the initial window (x0, x1, x2)
    bool go = elim(0);
    // rows 2 .. n-2 read (x[r-1], x[r], x[r+1]): shift in x[i+2] at step i
    int i = 1;
    if (go && n - 3 >= SPLINE_PF) {
        double px[SPLINE_PF], py[SPLINE_PF];
#pragma unroll
        for (int u = 0; u < SPLINE_PF; u++) {
            px[u] = XK(3 + u);
            py[u] = YK(3 + u);
        }
        for (; go && i + SPLINE_PF - 1 <= n - 3; i += SPLINE_PF) {
#pragma unroll
            for (int u = 0; u < SPLINE_PF; u++) {
                xa = xb; xb = xc; xc = px[u];
                ya = yb; yb = yc; yc = py[u];
                const int id = i + u + 2 + SPLINE_PF;
                if (id < n) {
                    px[u] = XK(id);
                    py[u] = YK(id);
                }
                if (!elim(i + u)) {
                    go = false;
                    break;
                }
            }
        }
    }
    for (; go && i <= n - 3; i++) {
        xa = xb; xb = xc; xc = XK(i + 2);
        ya = yb; yb = yc; yc = YK(i + 2);
        go = elim(i);
    }
    // the last row reads (x[n-3], x[n-2], x[n-1]), the window of row n-2
    if (go) go = elim(n - 2);
    if (info == 0 && cd == 0.0) {
        info = n;
        a.sB[(long)(n - 1) * ns + col] = cb;
        stop = n;
    }

    double *out = G.out;
    double c1L = 0.0, c2L = 0.0, c3L = 0.0;
    // Coefficients of interval i from slope(i) = xi and slope(i+1) = x_next
    // (set_spline_constants / fill_coefficients, interpolate.cu:28-43).
    auto emit = [&](int i, double xi, double x_next, double t0, double t1, double y0, double y1) {
        const double dt = t1 - t0;
        const double slope = (y1 - y0) / dt;
        const double tt = (xi + x_next - 2 * slope) / dt;
        const double c1 = xi;
        const double c2 = (slope - xi) / dt - tt;
        const double c3 = tt / dt;
        if (G.dense) {
            // cubic piece -> dense-output form (r5..r8 = 0), see few_oracle.c
            const double h = dt;
            const double h2 = h * h;
            const double h3 = h2 * h;
            const double q1 = c1 * h;
            const double q2 = c2 * h2;
            const double q3 = c3 * h3;
            const long no = (long)(n - 1) * G.nsys;
            const long o = (long)i * G.nsys + s;
            out[0 * no + o] = y0;
            out[1 * no + o] = (q1 + q2) + q3;
            out[2 * no + o] = -(q2 + q3);
            out[3 * no + o] = -q3;
            out[4 * no + o] = 0.0;
            out[5 * no + o] = 0.0;
            out[6 * no + o] = 0.0;
            out[7 * no + o] = 0.0;
        } else {
            const long o = (long)s * G.o_sys + (long)i * G.o_knot;
            out[o] = y0;
            out[G.o_coef + o] = c1;
            out[2 * G.o_coef + o] = c2;
            out[3 * G.o_coef + o] = c3;
        }
        if (i == n - 2) { c1L = c1; c2L = c2; c3L = c3; }
    };

    if (info != 0) {
        // LAPACK returns before back substitution; the reference then builds
        // coefficients from the partially eliminated right-hand side.  Rows
        // never reached keep their assembled values.
        if (a.info) *a.info = 1;  // benign race: every writer stores 1
        for (int r = stop; r < n; r++) {
            double b_, u_, d_, l_;
            int i0 = (r == 0) ? 0 : (r == n - 1 ? n - 3 : r - 1);
            prep_row(r, n, XK(i0), XK(i0 + 1), XK(i0 + 2), YK(i0), YK(i0 + 1), YK(i0 + 2), b_, u_, d_,
                     l_);
            a.sB[(long)r * ns + col] = b_;
        }
        double x_next = a.sB[(long)(n - 1) * ns + col];
        for (int ii = n - 2; ii >= 0; ii--) {
            const double xi = a.sB[(long)ii * ns + col];
            emit(ii, xi, x_next, XK(ii), XK(ii + 1), YK(ii), YK(ii + 1));
            x_next = xi;
        }
    } else {
        // ---------------- back substitution + coefficients ----------------
        // slope(i) is the solved knot derivative; interval i needs slope(i)
        // and slope(i+1): walk i downwards, emitting interval i as soon as
        // slope(i) is known.  Scratch rows and knots are prefetched
        // SPLINE_PF steps ahead.
        double x_next = cb / cd;  // slope(n-1)
        double x_next2 = 0.0;     // slope(i+2)
        double t1 = XK(n - 1), y1 = YK(n - 1);
        auto back = [&](int ii, double Bi, double DUi, double Di, double DLi, double t0, double y0) {
            double xi;
            if (ii == n - 2)
                xi = (Bi - DUi * x_next) / Di;
            else
                xi = (Bi - DUi * x_next - DLi * x_next2) / Di;
            emit(ii, xi, x_next, t0, t1, y0, y1);
            t1 = t0;
            y1 = y0;
            x_next2 = x_next;
            x_next = xi;
        };
        int ii = n - 2;
        if (n - 1 >= SPLINE_PF) {
            double rB[SPLINE_PF], rU[SPLINE_PF], rD[SPLINE_PF], rL[SPLINE_PF], rX[SPLINE_PF],
                rY[SPLINE_PF];
#pragma unroll
            for (int u = 0; u < SPLINE_PF; u++) {
                const long o = (long)(n - 2 - u) * ns + col;
                rB[u] = a.sB[o];
                rU[u] = a.sDU[o];
                rD[u] = a.sD[o];
                rL[u] = a.sDL[o];
                rX[u] = XK(n - 2 - u);
                rY[u] = YK(n - 2 - u);
            }
            for (; ii - (SPLINE_PF - 1) >= 0; ii -= SPLINE_PF) {
#pragma unroll
                for (int u = 0; u < SPLINE_PF; u++) {
                    const double Bi = rB[u], DUi = rU[u], Di = rD[u], DLi = rL[u];
                    const double t0 = rX[u], y0 = rY[u];
                    const int id = ii - u - SPLINE_PF;
                    if (id >= 0) {
                        const long o = (long)id * ns + col;
                        rB[u] = a.sB[o];
                        rU[u] = a.sDU[o];
                        rD[u] = a.sD[o];
                        rL[u] = a.sDL[o];
                        rX[u] = XK(id);
                        rY[u] = YK(id);
                    }
                    back(ii - u, Bi, DUi, Di, DLi, t0, y0);
                }
            }
        }
        for (; ii >= 0; ii--) {
            const long o = (long)ii * ns + col;
            back(ii, a.sB[o], a.sDU[o], a.sD[o], a.sDL[o], XK(ii), YK(ii));
        }
    }
    if (!G.dense) {
        // fill_final_derivs (interpolate.cu:242-280)
        const double x = XK(n - 1) - XK(n - 2);
        const double x2 = x * x;
        const long o = (long)s * G.o_sys + (long)(n - 1) * G.o_knot;
        out[o] = YK(n - 1);
        out[G.o_coef + o] = c1L + 2 * c2L * x + 3 * c3L * x2;
        out[2 * G.o_coef + o] = (2. * c2L + 6. * c3L * x) / 2.;
        out[3 * G.o_coef + o] = c3L;
    }
#undef XK
#undef YK
}

// Coefficients of interval i of system s from slope(i) = xi and slope(i+1) =
// x_next (set_spline_constants / fill_coefficients, interpolate.cu:28-43),
// written in the group's layout; the last interval's c1..c3 are returned for
// the final derivatives.
__host__ __device__ inline void spline_emit(const SplineGroup &G, int s, int n, int i, double xi,
                                            double x_next, double t0, double t1, double y0,
                                            double y1, double &c1L, double &c2L, double &c3L)
{
#pragma clang fp contract(off)
    double *out = G.out;
    const double dt = t1 - t0;
    const double slope = (y1 - y0) / dt;
    const double tt = (xi + x_next - 2 * slope) / dt;
    const double c1 = xi;
    const double c2 = (slope - xi) / dt - tt;
    const double c3 = tt / dt;
    if (G.dense) {
        const double h = dt;
        const double h2 = h * h;
        const double h3 = h2 * h;
        const double q1 = c1 * h;
        const double q2 = c2 * h2;
        const double q3 = c3 * h3;
        const long no = (long)(n - 1) * G.nsys;
        const long o = (long)i * G.nsys + s;
        out[0 * no + o] = y0;
        out[1 * no + o] = (q1 + q2) + q3;
        out[2 * no + o] = -(q2 + q3);
        out[3 * no + o] = -q3;
        out[4 * no + o] = 0.0;
        out[5 * no + o] = 0.0;
        out[6 * no + o] = 0.0;
        out[7 * no + o] = 0.0;
    } else {
        const long o = (long)s * G.o_sys + (long)i * G.o_knot;
        out[o] = y0;
        out[G.o_coef + o] = c1;
        out[2 * G.o_coef + o] = c2;
        out[3 * G.o_coef + o] = c3;
    }
    if (i == n - 2) {
        c1L = c1;
        c2L = c2;
        c3L = c3;
    }
}

// The matrix of the not-a-knot system on knots x (prep_row's d, du, dl with
// any y), eliminated as dgtsv does: F = piv | fact | d | du | dl (n each).
__host__ __device__ inline int spline_factor(const double *x, int n, double *F)
{
    double *piv = F, *fact = F + n, *d = F + 2 * n, *du = F + 3 * n, *dl = F + 4 * n;
    for (int r = 0; r < n; r++) {
        const int i0 = (r == 0) ? 0 : (r == n - 1 ? n - 3 : r - 1);
        double b_, u_, g_, l_;
        prep_row(r, n, x[i0], x[i0 + 1], x[i0 + 2], 0.0, 0.0, 0.0, b_, u_, g_, l_);
        d[r] = g_;
        if (r < n - 1) du[r] = u_;
        if (r > 0) dl[r - 1] = l_;
    }
    return dgtsv_factor(n, d, du, dl, piv, fact);
}

// One system on shared knots x with the recorded elimination F
// (spline_factor's piv | fact | d | du | dl), its right-hand side in the
// column bv (stride bs) as scratch:
//   forward: prep_row's right-hand side of row r = i + 1 formed on the fly
//            and swept with step i (dgtsv's no-interchange / interchange
//            update), one store per row;
//   back:    the back substitution, emitting interval ii's coefficients as
//            soon as its left slope is known, one load per row.
// The operations of spline_system_core in the same order for every value
// (bit-identical), without its per-system elimination.  The knot, value and
// elimination operands of the next PF steps are read ahead of the dependent
// chains (PF changes the schedule, not the arithmetic).
template <int PF>
__host__ __device__ inline void spline_system_shared(const SplineGroup &G, int s, int n,
                                                     const double *x, const double *F,
                                                     const double *yv, long ys, double *bv,
                                                     long bs)
{
#pragma clang fp contract(off)
#define SB(i) bv[(long)(i) * bs]
#define YV(i) yv[(long)(i) * ys]
    const double *piv = F, *fact = F + n, *d = F + 2 * n, *du = F + 3 * n, *dl = F + 4 * n;
    double xa = x[0], xb = x[1], xc = x[2];
    double ya = YV(0), yb = YV(1), yc = YV(2);
    double b_, u_, g_, l_;
    prep_row(0, n, xa, xb, xc, ya, yb, yc, b_, u_, g_, l_);
    double cur = b_;
    // step i with row r = i + 1 and its knot window
    auto step = [&](int i) {
        prep_row(i + 1, n, xa, xb, xc, ya, yb, yc, b_, u_, g_, l_);
        const bool keep = piv[i] == 0.0;
        SB(i) = keep ? cur : b_;
        cur = keep ? b_ - fact[i] * cur : cur - fact[i] * b_;
    };
    step(0);  // row 1 reads the window of row 0
    // rows 2 .. n-2 read (x[r-1], x[r], x[r+1]): shift in knot i + 2 at step i
    int i = 1;
    for (; i + PF - 1 <= n - 3; i += PF) {
        double px[PF], py[PF], pp[PF], pf[PF];
#pragma unroll
        for (int u = 0; u < PF; u++) {
            px[u] = x[i + u + 2];
            py[u] = YV(i + u + 2);
            pp[u] = piv[i + u];
            pf[u] = fact[i + u];
        }
#pragma unroll
        for (int u = 0; u < PF; u++) {
            xa = xb; xb = xc; xc = px[u];
            ya = yb; yb = yc; yc = py[u];
            prep_row(i + u + 1, n, xa, xb, xc, ya, yb, yc, b_, u_, g_, l_);
            const bool keep = pp[u] == 0.0;
            SB(i + u) = keep ? cur : b_;
            cur = keep ? b_ - pf[u] * cur : cur - pf[u] * b_;
        }
    }
    for (; i <= n - 3; i++) {
        xa = xb; xb = xc; xc = x[i + 2];
        ya = yb; yb = yc; yc = YV(i + 2);
        step(i);
    }
    step(n - 2);  // the last row reads the window of row n-2
    // back substitution + coefficients; slope(n-1) = cur / d[n-1]
    double x_next = cur / d[n - 1];
    double t1 = x[n - 1], y1 = YV(n - 1);
    const double yl = y1;
    double c1L = 0.0, c2L = 0.0, c3L = 0.0;
    // row n-2 has no second super-diagonal term
    double x_next2 = x_next;
    {
        const double t0 = x[n - 2], y0 = YV(n - 2);
        const double xi = (SB(n - 2) - du[n - 2] * x_next) / d[n - 2];
        spline_emit(G, s, n, n - 2, xi, x_next, t0, t1, y0, y1, c1L, c2L, c3L);
        t1 = t0;
        y1 = y0;
        x_next = xi;
    }
    int ii = n - 3;
    for (; ii - (PF - 1) >= 0; ii -= PF) {
        double rb[PF], ru[PF], rl[PF], rd[PF], rx[PF], ry[PF];
#pragma unroll
        for (int u = 0; u < PF; u++) {
            rb[u] = SB(ii - u);
            ru[u] = du[ii - u];
            rl[u] = dl[ii - u];
            rd[u] = d[ii - u];
            rx[u] = x[ii - u];
            ry[u] = YV(ii - u);
        }
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const double xi = (rb[u] - ru[u] * x_next - rl[u] * x_next2) / rd[u];
            spline_emit(G, s, n, ii - u, xi, x_next, rx[u], t1, ry[u], y1, c1L, c2L, c3L);
            t1 = rx[u];
            y1 = ry[u];
            x_next2 = x_next;
            x_next = xi;
        }
    }
    for (; ii >= 0; ii--) {
        const double t0 = x[ii], y0 = YV(ii);
        const double xi = (SB(ii) - du[ii] * x_next - dl[ii] * x_next2) / d[ii];
        spline_emit(G, s, n, ii, xi, x_next, t0, t1, y0, y1, c1L, c2L, c3L);
        t1 = t0;
        y1 = y0;
        x_next2 = x_next;
        x_next = xi;
    }
    if (!G.dense) {
        // fill_final_derivs (interpolate.cu:242-280)
        const double xx = x[n - 1] - x[n - 2];
        const double x2 = xx * xx;
        double *out = G.out;
        const long o = (long)s * G.o_sys + (long)(n - 1) * G.o_knot;
        out[o] = yl;
        out[G.o_coef + o] = c1L + 2 * c2L * xx + 3 * c3L * x2;
        out[2 * G.o_coef + o] = (2. * c2L + 6. * c3L * xx) / 2.;
        out[3 * G.o_coef + o] = c3L;
    }
#undef SB
#undef YV
}

__host__ __device__ void spline_one_system(const SplineArgs &a, int gi, int s)
{
    const SplineGroup &G = a.g[gi];
    if (s >= G.nsys) return;
    SplineScratch sc{a.sD, a.sDU, a.sDL, a.sB, a.info};
    spline_system_core(G, s, a.length, a.col_prefix[gi] + s, a.nsys_total, sc);
}

#define SPLINE_BLOCK 256  // threads of the shared-knot block (4 waves)
#define SPLINE_SPB 32     // systems per shared-knot block (half a wave)

// A block of SPLINE_SPB systems of a group on shared knots, four waves; the
// two halves of a wave take alternate knots / rows / intervals of the same
// systems.  Work that is parallel over (system, knot) is spread over the
// waves; only what is sequential along the knots runs on a single wave:
//   wave 3:     the shared matrix (prep_row's rows, one per lane), then
//               dgtsv's elimination chain (every lane the same chain);
//   waves 0-2:  meanwhile each stages the knot values of a third of the rows
//               (plus the neighbouring knots those rows read) and forms
//               their right-hand sides, reading only what it staged itself;
//   wave 0:     every system's sweeps (one lane per system) out of LDS;
//   waves 0-3:  the coefficients of the intervals and their stores.
__device__ void spline_block_shared(const SplineArgs &a, int gi, int blk, double *lds)
{
    constexpr int S = SPLINE_SPB;
    const SplineGroup &G = a.g[gi];
    const int n = a.length;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int half = lane / S, ls = lane % S;
    double *sx = lds, *F = sx + n, *sy = F + 5 * n, *sb = sy + (long)n * S;
    __shared__ int s_info;
    const int s0 = blk * S;
    const int nloc = min(S, G.nsys - s0);
    for (int i = tid; i < n; i += SPLINE_BLOCK) sx[i] = G.t[(long)i * G.t_knot];
    __syncthreads();
    if (wv == 3) {
        double *piv = F, *fact = F + n, *d = F + 2 * n, *du = F + 3 * n, *dl = F + 4 * n;
        for (int r = lane; r < n; r += 64) {
            const int i0 = (r == 0) ? 0 : (r == n - 1 ? n - 3 : r - 1);
            double b_, u_, g_, l_;
            prep_row(r, n, sx[i0], sx[i0 + 1], sx[i0 + 2], 0.0, 0.0, 0.0, b_, u_, g_, l_);
            d[r] = g_;
            if (r < n - 1) du[r] = u_;
            if (r > 0) dl[r - 1] = l_;
        }
        // the rows written by the other lanes of this wave, then the chain
        // (identical on every lane: identical stores, no divergence)
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const int info = dgtsv_factor_ahead<8>(n, d, du, dl, piv, fact);
        if (lane == 0) s_info = info;
    } else {
        const int c = (n + 2) / 3;
        const int r0 = wv * c, r1 = min(n, r0 + c);
        if (r0 < r1) {
            // knots [k0, k1] hold every value rows r0..r1-1 read; the two
            // halves stage alternate knots
            const int k0 = max(0, min(r0 - 1, n - 3)), k1 = min(n - 1, max(r1, 2));
            const double *ys = G.y + (long)(s0 + min(ls, nloc - 1)) * G.y_sys;
            for (int kb = k0 + half; kb <= k1; kb += 2 * SPLINE_TILE_BATCH) {
                double v[SPLINE_TILE_BATCH];
#pragma unroll
                for (int u = 0; u < SPLINE_TILE_BATCH; u++)
                    v[u] = ys[(long)min(kb + 2 * u, k1) * G.y_knot];
#pragma unroll
                for (int u = 0; u < SPLINE_TILE_BATCH; u++)
                    if (kb + 2 * u <= k1) sy[(long)(kb + 2 * u) * S + ls] = ls < nloc ? v[u] : 0.0;
            }
            // the other half's knots, written by this wave's earlier LDS stores
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            // right-hand sides (prep_row's b), alternate rows per half
#pragma unroll 2
            for (int r = r0 + half; r < r1; r += 2) {
                const int i0 = (r == 0) ? 0 : (r == n - 1 ? n - 3 : r - 1);
                double b_, u_, g_, l_;
                prep_row(r, n, sx[i0], sx[i0 + 1], sx[i0 + 2], sy[(long)i0 * S + ls],
                         sy[(long)(i0 + 1) * S + ls], sy[(long)(i0 + 2) * S + ls], b_, u_, g_, l_);
                sb[(long)r * S + ls] = b_;
            }
        }
    }
    __syncthreads();
    if (s_info != 0) {  // singular: dgtsv's early return, handled system by system
        if (wv == 0 && half == 0) spline_one_system(a, gi, s0 + ls);
        return;
    }
    if (wv == 0 && half == 0) dgtsv_sweeps(n, F, sb + ls, S);
    __syncthreads();
    if (ls >= nloc) return;
    // coefficients of interval i on lane group i % 8 (wave, half); the group
    // holding interval n - 2 also writes the final derivatives
    // (fill_final_derivs, interpolate.cu:242-280)
    const int grp = 2 * wv + half;
    const int s = s0 + ls;
    double c1L = 0.0, c2L = 0.0, c3L = 0.0;
#pragma unroll 2
    for (int i = grp; i < n - 1; i += 8)
        spline_emit(G, s, n, i, sb[(long)i * S + ls], sb[(long)(i + 1) * S + ls], sx[i], sx[i + 1],
                    sy[(long)i * S + ls], sy[(long)(i + 1) * S + ls], c1L, c2L, c3L);
    if (!G.dense && grp == (n - 2) % 8) {
        const double xx = sx[n - 1] - sx[n - 2];
        const double x2 = xx * xx;
        double *out = G.out;
        const long o = (long)s * G.o_sys + (long)(n - 1) * G.o_knot;
        out[o] = sy[(long)(n - 1) * S + ls];
        out[G.o_coef + o] = c1L + 2 * c2L * xx + 3 * c3L * x2;
        out[2 * G.o_coef + o] = (2. * c2L + 6. * c3L * xx) / 2.;
        out[3 * G.o_coef + o] = c3L;
    }
}

__global__ __launch_bounds__(SPLINE_BLOCK) void spline_build_kernel(SplineArgs a)
{
    extern __shared__ __attribute__((aligned(16))) double spl_lds[];
    int gi = 0;
    while (gi + 1 < a.ngroups && (int)blockIdx.x >= a.blk_prefix[gi + 1]) gi++;
    const int blk = (int)blockIdx.x - a.blk_prefix[gi];
    if (a.shared_lds && a.g[gi].t_sys == 0) {  // uniform over the block
        spline_block_shared(a, gi, blk, spl_lds);
        return;
    }
    if (threadIdx.x < 64) spline_one_system(a, gi, blk * 64 + (int)threadIdx.x);
}

void fill_spline_args(const SplineGroup *groups, int ngroups, int length, double *scratch4[4],
                      int *info_out, SplineArgs &a)
{
    a.ngroups = ngroups;
    a.length = length;
    a.blk_prefix[0] = 0;
    a.col_prefix[0] = 0;
    for (int g = 0; g < SPLINE_MAX_GROUPS; g++) {
        a.g[g] = groups[g < ngroups ? g : 0];
        const int ns = g < ngroups ? groups[g].nsys : 0;
        a.blk_prefix[g + 1] = a.blk_prefix[g] + (ns + 63) / 64;
        a.col_prefix[g + 1] = a.col_prefix[g] + ns;
    }
    a.nsys_total = a.col_prefix[ngroups];
    a.sD = scratch4[0];
    a.sDU = scratch4[1];
    a.sDL = scratch4[2];
    a.sB = scratch4[3];
    a.info = info_out;
    a.shared_lds = 0;
}

// Batched build: every system of every waveform (2K amplitude arrays, then
// the 3 phase arrays in dense-output form unless the phases come as dense
// output already, p.pk == nullptr).  All systems of a waveform share its knot
// matrix, so it is eliminated once per waveform (batch_spline_factor_kernel)
// and each system (one lane) only sweeps its right-hand side with the
// recorded steps: one scratch plane per system instead of four, and no
// division of the elimination per system.

// One 64-lane block per waveform: the matrix rows (prep_row with y = 0, one
// lane per row), then dgtsv's elimination on lane 0; INFO after the five
// arrays.
__global__ __launch_bounds__(64) void batch_spline_factor_kernel(BatchDev B, BatchPtrs p)
{
    const WaveDesc d = B.desc[blockIdx.x];
    const int n = d.L;
    const double *x = p.t + d.t_off;
    double *F = p.scratch + d.sc_off;
    double *piv = F, *fact = F + n, *dg = F + 2 * n, *du = F + 3 * n, *dl = F + 4 * n;
    for (int r = threadIdx.x; r < n; r += 64) {
        const int i0 = (r == 0) ? 0 : (r == n - 1 ? n - 3 : r - 1);
        double b_, u_, g_, l_;
        prep_row(r, n, x[i0], x[i0 + 1], x[i0 + 2], 0.0, 0.0, 0.0, b_, u_, g_, l_);
        dg[r] = g_;
        if (r < n - 1) du[r] = u_;
        if (r > 0) dl[r - 1] = l_;
    }
    __syncthreads();  // the rows of every lane, for lane 0's chain
    if (threadIdx.x == 0) {
        const int info = dgtsv_factor_ahead<8>(n, dg, du, dl, piv, fact);
        *reinterpret_cast<int *>(F + 5 * n) = info;
    }
}

// System sg of the batch: its group (amplitude real / imaginary parts or
// phases), index in the group and waveform.  False for phase systems when
// the phases come as dense output (nothing to build).
struct BatchSystem {
    SplineGroup G;
    int s;
    long ls;
    WaveDesc d;
};

__device__ inline bool batch_system(const BatchDev &B, const BatchPtrs &p, long sg, BatchSystem &S)
{
    const int b = upper_index(B.sys_prefix, B.B, sg);
    S.d = B.desc[b];
    S.ls = sg - B.sys_prefix[b];
    const long K = S.d.K, L = S.d.L;
    const double *teuk = p.teuk + S.d.teuk_off;
    const double *t = p.t + S.d.t_off;
    double *interp = p.interp + S.d.interp_off;
    if (S.ls < K) {
        S.G = SplineGroup{teuk, 2, 2 * K, t, 0, 1, interp, L * 2 * K, 1, 2 * K, (int)K, 0};
        S.s = (int)S.ls;
    } else if (S.ls < 2 * K) {
        S.G = SplineGroup{teuk + 1, 2, 2 * K, t, 0, 1, interp + K, L * 2 * K, 1, 2 * K, (int)K, 0};
        S.s = (int)(S.ls - K);
    } else {
        if (!p.pk) return false;
        S.G = SplineGroup{p.pk + S.d.pk_off, L, 1, t, 0, 1, p.pcoef + S.d.pc_off, 0, 0, 0, 3, 1};
        S.s = (int)(S.ls - 2 * K);
    }
    return true;
}

// One lane per system of a waveform whose matrix was eliminated (INFO = 0).
__global__ __launch_bounds__(64) void batch_spline_kernel(BatchDev B, BatchPtrs p)
{
    const long sg = (long)blockIdx.x * 64 + threadIdx.x;
    BatchSystem S;
    if (sg >= B.total_systems || !batch_system(B, p, sg, S)) return;
    const long L = S.d.L;
    const double *F = p.scratch + S.d.sc_off;
    if (*reinterpret_cast<const int *>(F + 5 * L) != 0) return;
    double *plane = p.scratch + S.d.sc_off + batch_spline_head(L);
    spline_system_shared<2>(S.G, S.s, (int)L, S.G.t, F, S.G.y + (long)S.s * S.G.y_sys, S.G.y_knot,
                            plane + S.ls, 2 * S.d.K + 3);
}

// The systems of waveforms whose matrix is singular: LAPACK returns early
// at a step that the right-hand side's values decide, so each system runs
// dgtsv on its own (a separate kernel keeps batch_spline_kernel's register
// budget that of the swept path).
__global__ __launch_bounds__(64) void batch_spline_single_kernel(BatchDev B, BatchPtrs p)
{
    const long sg = (long)blockIdx.x * 64 + threadIdx.x;
    BatchSystem S;
    if (sg >= B.total_systems || !batch_system(B, p, sg, S)) return;
    const long L = S.d.L, ns = 2 * S.d.K + 3;
    if (*reinterpret_cast<const int *>(p.scratch + S.d.sc_off + 5 * L) == 0) return;
    double *planes = p.scratch + S.d.sc_off + batch_spline_head(L);
    const long f = L * ns;
    SplineScratch scr{planes, planes + f, planes + 2 * f, planes + 3 * f, nullptr};
    spline_system_core(S.G, S.s, (int)L, S.ls, ns, scr);
}

int launch_batch_splines(const BatchDev &B, const BatchPtrs &p, hipStream_t stream)
{
    if (B.total_systems <= 0) return FEW_OK;
    const long blocks = (B.total_systems + 63) / 64;
    hipLaunchKernelGGL(batch_spline_factor_kernel, dim3((unsigned)B.B), dim3(64), 0, stream, B, p);
    hipLaunchKernelGGL(batch_spline_kernel, dim3((unsigned)blocks), dim3(64), 0, stream, B, p);
    hipLaunchKernelGGL(batch_spline_single_kernel, dim3((unsigned)blocks), dim3(64), 0, stream, B, p);
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

int launch_spline_build(const SplineGroup *groups, int ngroups, int length, double *scratch4[4],
                        int *info_out, hipStream_t stream)
{
    if (ngroups < 1 || ngroups > SPLINE_MAX_GROUPS || length < 3) {
        set_error("spline build: need 1-%d groups and length >= 3 (got %d groups, length %d)",
                  SPLINE_MAX_GROUPS, ngroups, length);
        return FEW_ERR_ARG;
    }
    SplineArgs a;
    fill_spline_args(groups, ngroups, length, scratch4, info_out, a);
    if (a.blk_prefix[ngroups] == 0) return FEW_OK;
    size_t lds = 0;
    if (length <= SPLINE_SHARED_MAX_L) {
        for (int g = 0; g < ngroups; g++)
            if (groups[g].t_sys == 0) a.shared_lds = 1;
        if (a.shared_lds) {
            lds = sizeof(double) * (size_t)length * (SPLINE_SPB * 2 + 6);
            FEW_LDS(spline_build_kernel, (int)lds);
            // groups on shared knots take SPLINE_SPB systems per block
            for (int g = 0; g < ngroups; g++) {
                const int per = groups[g].t_sys == 0 ? SPLINE_SPB : 64;
                a.blk_prefix[g + 1] = a.blk_prefix[g] + (groups[g].nsys + per - 1) / per;
            }
        }
    }
    const int nblk = a.blk_prefix[ngroups];
    hipLaunchKernelGGL(spline_build_kernel, dim3(nblk), dim3(SPLINE_BLOCK), lds, stream, a);
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

}  // namespace few

// Frequency-domain interpolated sum: the stationary-phase approximation of
// every (mode, knot segment) evaluated at every frequency bin the segment's
// mode frequency sweeps over (get_waveform_generic_fd,
// src/few/cutils/interpolate.cu:1142-1563).
//
// Per bin f of item (mode q, segment j) -- interpolate.cu:1396-1500:
//   * the times where the mode frequency F(x) = m f_phi + k f_theta + n f_r
//     (cubic in x = t - t_j) equals f: Cardano's roots of
//     F3 x^3 + F2 x^2 + F1 x + (F0 - f), keeping the real part of each of the
//     three complex roots and every one that falls inside the segment;
//   * at such a time t: amplitude A(x), phase Phi(t) and its first two time
//     derivatives from the dense-output form, and
//       h(f) += Y+ A T(fdot, fddot) e^{i(2 pi f t - Phi)},
//       h(-f) += [m != 0] Y- conj(A T e^{...})      (bin 2 zero - i)
//     with T = -fdot/|fddot| (2/sqrt 3) SPA(x) / sqrt(x),
//     x = 2 pi fdot^3 / (3 fddot^2) and SPA(x) = i sqrt(x) e^{-ix} K_1/3(-ix).
//
// Layout of the work (summed output): bin-owned.  Each thread owns a bin and
// its mirror about the zero bin, a workgroup a tile of such owners; the
// segments that reach the tile are found per mode by binary search over the
// mode's segments sorted by first bin, and each thread adds every
// contribution to its bins itself, in a fixed order (mode, sorted segment,
// root): no atomics, and the result is the same on every run (the
// reference's kernel adds with atomic CAS loops in no fixed order).  With
// separate_modes every (mode, bin) has one writer already: the (mode,
// segment) items are flattened, their in-range bin counts prefix-summed on
// the device, and a persistent grid walks the flattened bins.
//
// Complex arithmetic follows gcmplx::complex (src/few/cutils/cuda_complex.hpp,
// the libc++ <complex> algorithms) operation for operation, on the device's
// FP64 libm, with contraction off in the whole file; the products keep their
// C99 Annex G NaN recovery.  cos and sin of one argument are taken with one
// sincos call, as an optimising compiler emits for gcmplx::exp / polar.
#include <math.h>

#include <algorithm>

#include "fd.hpp"
#include "few_common.hpp"

#pragma clang fp contract(off)

namespace few {

struct fcx {
    double re, im;
};

__host__ __device__ inline fcx fcx_make(double re, double im)
{
    fcx z;
    z.re = re;
    z.im = im;
    return z;
}

// cuda_complex.hpp:440-498
__host__ __device__ inline fcx fcx_mul(fcx z, fcx w)
{
#pragma clang fp contract(off)
    double a = z.re, b = z.im, c = w.re, d = w.im;
    const double ac = a * c, bd = b * d, ad = a * d, bc = b * c;
    double x = ac - bd, y = ad + bc;
    if (isnan(x) && isnan(y)) {
        bool recalc = false;
        if (isinf(a) || isinf(b)) {
            a = copysign(isinf(a) ? 1.0 : 0.0, a);
            b = copysign(isinf(b) ? 1.0 : 0.0, b);
            if (isnan(c)) c = copysign(0.0, c);
            if (isnan(d)) d = copysign(0.0, d);
            recalc = true;
        }
        if (isinf(c) || isinf(d)) {
            c = copysign(isinf(c) ? 1.0 : 0.0, c);
            d = copysign(isinf(d) ? 1.0 : 0.0, d);
            if (isnan(a)) a = copysign(0.0, a);
            if (isnan(b)) b = copysign(0.0, b);
            recalc = true;
        }
        if (!recalc && (isinf(ac) || isinf(bd) || isinf(ad) || isinf(bc))) {
            if (isnan(a)) a = copysign(0.0, a);
            if (isnan(b)) b = copysign(0.0, b);
            if (isnan(c)) c = copysign(0.0, c);
            if (isnan(d)) d = copysign(0.0, d);
            recalc = true;
        }
        if (recalc) {
            x = INFINITY * (a * c - b * d);
            y = INFINITY * (a * d + b * c);
        }
    }
    return fcx_make(x, y);
}

// complex (*, /) real component-wise; real + complex adds to the real part
__host__ __device__ inline fcx fcx_scale(fcx z, double s) { return fcx_make(z.re * s, z.im * s); }
__host__ __device__ inline fcx fcx_divr(fcx z, double s) { return fcx_make(z.re / s, z.im / s); }
__host__ __device__ inline fcx fcx_add(fcx z, fcx w) { return fcx_make(z.re + w.re, z.im + w.im); }
__host__ __device__ inline fcx fcx_addr(double r, fcx z) { return fcx_make(z.re + r, z.im); }

// cuda_complex.hpp operator/(complex, complex): logb/scalbn-scaled quotient
__host__ __device__ inline fcx fcx_div(fcx z, fcx w)
{
#pragma clang fp contract(off)
    int ilogbw = 0;
    double a = z.re, b = z.im, c = w.re, d = w.im;
    const double logbw = logb(fmax(fabs(c), fabs(d)));
    if (isfinite(logbw)) {
        ilogbw = (int)logbw;
        c = scalbn(c, -ilogbw);
        d = scalbn(d, -ilogbw);
    }
    const double denom = c * c + d * d;
    double x = scalbn((a * c + b * d) / denom, -ilogbw);
    double y = scalbn((b * c - a * d) / denom, -ilogbw);
    if (isnan(x) && isnan(y)) {
        if ((denom == 0.0) && (!isnan(a) || !isnan(b))) {
            x = copysign(INFINITY, c) * a;
            y = copysign(INFINITY, c) * b;
        } else if ((isinf(a) || isinf(b)) && isfinite(c) && isfinite(d)) {
            a = copysign(isinf(a) ? 1.0 : 0.0, a);
            b = copysign(isinf(b) ? 1.0 : 0.0, b);
            x = INFINITY * (a * c + b * d);
            y = INFINITY * (b * c - a * d);
        } else if (isinf(logbw) && logbw > 0.0 && isfinite(a) && isfinite(b)) {
            c = copysign(isinf(c) ? 1.0 : 0.0, c);
            d = copysign(isinf(d) ? 1.0 : 0.0, d);
            x = 0.0 * (a * c + b * d);
            y = 0.0 * (b * c - a * d);
        }
    }
    return fcx_make(x, y);
}

__host__ __device__ inline fcx fcx_rdiv(double x, fcx w) { return fcx_div(fcx_make(x, 0.0), w); }
__host__ __device__ inline double fcx_abs(fcx z) { return hypot(z.re, z.im); }
__host__ __device__ inline double fcx_arg(fcx z) { return atan2(z.im, z.re); }

// cuda_complex.hpp:864-893
__host__ __device__ inline fcx fcx_polar(double rho, double theta)
{
#pragma clang fp contract(off)
    if (isnan(rho) || signbit(rho)) return fcx_make(NAN, NAN);
    if (isnan(theta)) {
        if (isinf(rho)) return fcx_make(rho, theta);
        return fcx_make(theta, theta);
    }
    if (isinf(theta)) {
        if (isinf(rho)) return fcx_make(rho, NAN);
        return fcx_make(NAN, NAN);
    }
    double sn, cs;
    sincos(theta, &sn, &cs);
    double x = rho * cs;
    if (isnan(x)) x = 0;
    double y = rho * sn;
    if (isnan(y)) y = 0;
    return fcx_make(x, y);
}

__host__ __device__ inline fcx fcx_log(fcx z) { return fcx_make(log(fcx_abs(z)), fcx_arg(z)); }

// cuda_complex.hpp:914-930
__host__ __device__ inline fcx fcx_sqrt(fcx z)
{
    if (isinf(z.im)) return fcx_make(INFINITY, z.im);
    if (isinf(z.re)) {
        if (z.re > 0.0) return fcx_make(z.re, isnan(z.im) ? z.im : copysign(0.0, z.im));
        return fcx_make(isnan(z.im) ? z.im : 0.0, copysign(z.re, z.im));
    }
    return fcx_polar(sqrt(fcx_abs(z)), fcx_arg(z) / 2.0);
}

// cuda_complex.hpp:934-958
__host__ __device__ inline fcx fcx_exp(fcx z)
{
#pragma clang fp contract(off)
    double i = z.im;
    if (isinf(z.re)) {
        if (z.re < 0.0) {
            if (!isfinite(i)) i = 1.0;
        } else if (i == 0 || !isfinite(i)) {
            if (isinf(i)) i = NAN;
            return fcx_make(z.re, i);
        }
    } else if (isnan(z.re) && z.im == 0) {
        return z;
    }
    const double e = exp(z.re);
    double sn, cs;
    sincos(i, &sn, &cs);
    return fcx_make(e * cs, e * sn);
}

// cuda_complex.hpp:960-976: pow(x, y) = exp(complex(y) * log(x))
__host__ __device__ inline fcx fcx_pow(fcx x, double y)
{
    return fcx_exp(fcx_mul(fcx_make(y, 0.0), fcx_log(x)));
}

// The argument-independent factors of SPAFunc: exp(-2 pi i / 3), exp(-pi i / 3)
// and exp(-3 pi i / 4) sqrt(pi / 2), computed once per thread with the same
// operations the reference repeats on every call.
struct FdSpaConst {
    fcx e1, e2, pref_asym;
};

__host__ __device__ inline FdSpaConst fd_spa_const()
{
    const fcx II = fcx_make(0.0, 1.0);
    FdSpaConst k;
    k.e1 = fcx_exp(fcx_divr(fcx_scale(II, -2. * M_PI), 3.));
    k.e2 = fcx_exp(fcx_divr(fcx_scale(II, -M_PI), 3.));
    k.pref_asym = fcx_scale(fcx_exp(fcx_scale(fcx_scale(II, -0.75), M_PI)), sqrt(0.5 * M_PI));
    return k;
}

// SPAFunc (interpolate.cu:1041-1090): i sqrt(x) e^{-ix} K_1/3(-ix) by its
// small-argument series (|x| <= 7) or its asymptotic series.
__host__ __device__ inline fcx fd_spa(double x, const FdSpaConst &kc)
{
#pragma clang fp contract(off)
    const fcx II = fcx_make(0.0, 1.0);
    const double Gamp13 = 2.67893853470774763;   // Gamma(1/3)
    const double Gamm13 = -4.06235381827920125;  // Gamma(-1/3)
    if (fabs(x) <= 7.) {
        const fcx tmp_p = fcx_pow(fcx_make(x, 0.0), 1. / 6.);
        fcx pref1 = fcx_mul(kc.e1, tmp_p);
        pref1 = fcx_mul(pref1, tmp_p);
        pref1 = fcx_mul(pref1, tmp_p);
        pref1 = fcx_mul(pref1, tmp_p);
        pref1 = fcx_mul(pref1, tmp_p);
        pref1 = fcx_divr(fcx_scale(pref1, Gamm13), 1.2599210498948732);  // 2^(1/3)
        fcx pref2 = fcx_mul(kc.e2, tmp_p);
        pref2 = fcx_divr(fcx_scale(pref2, Gamp13), 1.5874010519681994748);  // 2^(2/3)
        const double x2 = x * x;
        // Taylor coefficients of the two Bessel-series factors (even powers)
        const double ser1 = 0.5 + x2 * (-0.09375 + x2 * (0.0050223214285714285714 + x2 * (-0.00012555803571428571429 + x2 * (1.8109332074175824176e-6 + x2 * (-1.6977498819539835165e-8 + x2 * (1.1169407118118312608e-10 + x2 * (-5.4396463237589184781e-13 + x2 * (2.0398673714095944293e-15 + x2 * (-6.0710338434809358015e-18 + x2 * (1.4687985105195812423e-20 + x2 * (-2.9454515585285720100e-23 + x2 * (4.9754249299469121790e-26 + x2 * -7.1760936489618925658e-29))))))))))));
        const double ser2 = 1. + x2 * (-0.375 + x2 * (0.028125 + x2 * (-0.00087890625 + x2 * (0.000014981356534090909091 + x2 * (-1.6051453429383116883e-7 + x2 * (1.1802539286311115355e-9 + x2 * (-6.3227889033809546546e-12 + x2 * (2.5772237377911499951e-14 + x2 * (-8.2603324929203525483e-17 + x2 * (2.1362928861000911763e-19 + x2 * (-4.5517604107246260858e-22 + x2 * (8.1281435905796894390e-25 + x2 * -1.2340298973552160079e-27))))))))))));
        const fcx ex = fcx_exp(fcx_scale(fcx_make(-II.re, -II.im), x));
        return fcx_mul(ex, fcx_add(fcx_scale(pref1, ser1), fcx_scale(pref2, ser2)));
    }
    const fcx y = fcx_make(1. / x, 0.0);
    const fcx pref = kc.pref_asym;
    // asymptotic series c_0 + y (c_1 + ... + y c_8): alternating imaginary/real terms
    fcx ser = fcx_add(fcx_make(-0.87766696951001691647, 0.0), fcx_mul(y, fcx_scale(II, 3.0794530301731669934)));
    ser = fcx_add(fcx_scale(II, -0.29159139923075051147), fcx_mul(y, ser));
    ser = fcx_add(fcx_make(0.11609906402551541102, 0.0), fcx_mul(y, ser));
    ser = fcx_add(fcx_scale(II, 0.057649190412669721333), fcx_mul(y, ser));
    ser = fcx_add(fcx_make(-0.037993059127800640146, 0.0), fcx_mul(y, ser));
    ser = fcx_add(fcx_scale(II, -0.037133487654320987654), fcx_mul(y, ser));
    ser = fcx_add(fcx_make(0.069444444444444444444, 0.0), fcx_mul(y, ser));
    ser = fcx_add(II, fcx_mul(y, ser));
    return fcx_mul(pref, ser);
}

// cube_roots (interpolate.cu:1113-1134): Cardano with the principal complex
// cube root; xi = (-1 + sqrt(-3)) / 2 is passed in (computed once per thread).
__host__ __device__ inline void fd_cube_roots(double a, double b, double c, double d, fcx xi,
                                              double out[3])
{
#pragma clang fp contract(off)
    const double b2 = b * b;
    const double b3 = b2 * b;
    const double a2 = a * a;
    const double D0 = (b2 - 3 * a * c);
    const double D1 = (2 * b3 - 9 * a * b * c + 27 * a2 * d);
    const fcx sq = fcx_sqrt(fcx_make(D1 * D1 - 4 * (D0 * D0 * D0), 0.0));
    const fcx C = fcx_pow(fcx_divr(fcx_addr(D1, sq), 2.), 1. / 3.);
    const double pre = -(1. / (3 * a));
    const fcx xiC = fcx_mul(xi, C);
    const fcx xixiC = fcx_mul(fcx_mul(xi, xi), C);
    out[0] = fcx_scale(fcx_add(fcx_addr(b, C), fcx_rdiv(D0, C)), pre).re;
    out[1] = fcx_scale(fcx_add(fcx_addr(b, xiC), fcx_rdiv(D0, xiC)), pre).re;
    out[2] = fcx_scale(fcx_add(fcx_addr(b, xixiC), fcx_rdiv(D0, xixiC)), pre).re;
}

__host__ __device__ inline fcx fd_xi()
{
    return fcx_divr(fcx_addr(-1., fcx_sqrt(fcx_make(-3., 0.0))), 2.);
}

// Everything one (mode, segment) item reads (interpolate.cu:1240-1345).
struct FdItem {
    double ar[4], ai[4];          // amplitude cubic, Re and Im
    double pp[8], pt[8], pr[8];   // dense-output phases
    double F0, F1, F2, F3;        // mode-frequency cubic
    double t0, t1, pt0, pw;       // segment bounds, phase segment start and width
    double mm, kk, nn;
    fcx yp, ym;
};

__host__ __device__ inline void fd_load_item(const FdArgs &a, int q, int j, FdItem &it)
{
#pragma clang fp contract(off)
    const long nb = (2L * a.K + 3) * a.L;
    const long np = (long)(a.L - 1) * 3;
    double fph[4], fth[4], frr[4];
    for (int c = 0; c < 4; c++) {
        it.ar[c] = a.interp[c * nb + (long)q * a.L + j];
        it.ai[c] = a.interp[c * nb + (long)(a.K + q) * a.L + j];
        fph[c] = a.interp[c * nb + (long)(2 * a.K) * a.L + j];
        fth[c] = a.interp[c * nb + (long)(2 * a.K + 1) * a.L + j];
        frr[c] = a.interp[c * nb + (long)(2 * a.K + 2) * a.L + j];
    }
    for (int c = 0; c < 8; c++) {
        it.pp[c] = a.pcoef[c * np + j];
        it.pt[c] = a.pcoef[c * np + (a.L - 1) + j];
        it.pr[c] = a.pcoef[c * np + 2 * (a.L - 1) + j];
    }
    const int m = a.m[q], k = a.k[q], n = a.n[q];
    it.mm = m;
    it.kk = k;
    it.nn = n;
    it.F0 = m * fph[0] + k * fth[0] + n * frr[0];
    it.F1 = m * fph[1] + k * fth[1] + n * frr[1];
    it.F2 = m * fph[2] + k * fth[2] + n * frr[2];
    it.F3 = m * fph[3] + k * fth[3] + n * frr[3];
    it.t0 = a.t_knots[j];
    it.t1 = a.t_knots[j + 1];
    it.pt0 = a.phase_t[j];
    it.pw = a.phase_t[j + 1] - a.phase_t[j];
    it.yp = fcx_make(a.ylm[q].x, a.ylm[q].y);
    it.ym = fcx_make(a.ylm[a.K + q].x, a.ylm[a.K + q].y);
}

__host__ __device__ inline double fd_phase(const double *r, double s, double s1)
{
#pragma clang fp contract(off)
    return r[0] + s * (r[1] + s1 * (r[2] + s * (r[3] + s1 * (r[4] + s * (r[5] + s1 * (r[6] + s * r[7]))))));
}

__host__ __device__ inline double fd_fdot(const double *r, double s, double s2, double s3,
                                          double s4, double s5, double w2)
{
#pragma clang fp contract(off)
    return (-2. * r[2] + r[3] * (2. - 6. * s) + r[4] * (2. - 12. * s + 12. * s2) +
            r[5] * (6. * s - 24. * s2 + 20. * s3) + r[6] * (6. * s - 36. * s2 + 60. * s3 - 30. * s4) +
            r[7] * (12. * s2 - 60. * s3 + 90. * s4 - 42. * s5)) /
           (2. * M_PI * w2);
}

__host__ __device__ inline double fd_fddot(const double *r, double s, double s2, double s3,
                                           double s4, double w3)
{
#pragma clang fp contract(off)
    return (-6. * r[3] + r[4] * (-12. + 24. * s) + r[5] * (6. - 48. * s + 60. * s2) +
            r[6] * (6. - 72. * s + 180. * s2 - 120. * s3) +
            r[7] * (24. * s - 180. * s2 + 360. * s3 - 210. * s4)) /
           (2. * M_PI * w3);
}

// The contributions of bin i of an item: one per root inside the segment
// (up to three), each the +f term and the -f (conjugate) term, handed to
// emit(plus, minus) in root order.
template <class Emit>
__host__ __device__ inline void fd_bin(const FdItem &it, double f, fcx xi, const FdSpaConst &kc,
                                       Emit emit)
{
#pragma clang fp contract(off)
    double roots[3];
    fd_cube_roots(it.F3, it.F2, it.F1, (it.F0 - f), xi, roots);
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const double root = roots[r];
        const double t = it.t0 + root;
        if (!((t < it.t1) && (t >= it.t0))) continue;
        const double x = root, x2 = x * x, x3 = x * x2;
        const double Are = it.ar[0] + it.ar[1] * x + it.ar[2] * x2 + it.ar[3] * x3;
        const double Aim = it.ai[0] + it.ai[1] * x + it.ai[2] * x2 + it.ai[3] * x3;
        const double s = (t - it.pt0) / it.pw;
        const double s1 = 1.0 - s;
        const double s2 = s * s, s3 = s * s2, s4 = s * s3, s5 = s * s4;
        const double w2 = it.pw * it.pw;
        const double w3 = it.pw * w2;
        const double phase = it.mm * fd_phase(it.pp, s, s1) + it.kk * fd_phase(it.pt, s, s1) +
                             it.nn * fd_phase(it.pr, s, s1);
        const double fdot = it.mm * fd_fdot(it.pp, s, s2, s3, s4, s5, w2) +
                            it.kk * fd_fdot(it.pt, s, s2, s3, s4, s5, w2) +
                            it.nn * fd_fdot(it.pr, s, s2, s3, s4, s5, w2);
        const double fddot = it.mm * fd_fddot(it.pp, s, s2, s3, s4, w3) +
                             it.kk * fd_fddot(it.pt, s, s2, s3, s4, w3) +
                             it.nn * fd_fddot(it.pr, s, s2, s3, s4, w3);
        const double targ = 2. * M_PI * fdot * fdot * fdot / (3. * fddot * fddot);
        const double pre = -1.0 * fdot / fabs(fddot) * 2. / sqrt(3.);
        const fcx amp2 = fcx_div(fcx_scale(fd_spa(targ, kc), pre), fcx_sqrt(fcx_make(targ, 0.0)));
        const fcx ph = fcx_exp(fcx_scale(fcx_make(0.0, 1.0), 2. * M_PI * f * t - phase));
        const fcx te = fcx_mul(fcx_mul(fcx_make(Are, Aim), amp2), ph);
        emit(fcx_mul(it.yp, te), fcx_mul(it.ym, fcx_make(te.re, -te.im)));
    }
}

__host__ __device__ inline int fd_minus_index(int zero_index, int i)
{
    const int diff = abs(zero_index - i);
    return (i < zero_index) ? zero_index + diff : zero_index - diff;
}

// In-range bins of item (q, j): [max(start, 0), min(end, Nf - 1)].
__host__ __device__ inline void fd_item_range(const FdArgs &a, long item, int &lo, int &hi)
{
    lo = a.seg_start[item];
    hi = a.seg_end[item];
    if (lo < 0) lo = 0;
    if (hi > a.data_length - 1) hi = a.data_length - 1;
}

#define FD_SCAN_BLOCK 1024

// Block-local exclusive scan of the per-item bin counts.
__global__ __launch_bounds__(FD_SCAN_BLOCK) void fd_scan1_kernel(FdArgs a, long nitems, long *local,
                                                                 long *bsum)
{
    __shared__ long s[FD_SCAN_BLOCK];
    const int tid = threadIdx.x;
    const long q = (long)blockIdx.x * FD_SCAN_BLOCK + tid;
    long c = 0;
    if (q < nitems) {
        int lo, hi;
        fd_item_range(a, q, lo, hi);
        c = hi >= lo ? (long)(hi - lo + 1) : 0;
    }
    s[tid] = c;
    __syncthreads();
    for (int off = 1; off < FD_SCAN_BLOCK; off <<= 1) {
        const long v = tid >= off ? s[tid - off] : 0;
        __syncthreads();
        s[tid] += v;
        __syncthreads();
    }
    if (q < nitems) local[q] = s[tid] - c;
    if (tid == FD_SCAN_BLOCK - 1) bsum[blockIdx.x] = s[tid];
}

// Exclusive scan of the block sums (one workgroup, sequential over chunks).
__global__ __launch_bounds__(FD_SCAN_BLOCK) void fd_scan2_kernel(long *bsum, long nb, long *total)
{
    __shared__ long s[FD_SCAN_BLOCK];
    __shared__ long carry;
    const int tid = threadIdx.x;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (long b0 = 0; b0 < nb; b0 += FD_SCAN_BLOCK) {
        const long b = b0 + tid;
        const long v0 = b < nb ? bsum[b] : 0;
        s[tid] = v0;
        __syncthreads();
        for (int off = 1; off < FD_SCAN_BLOCK; off <<= 1) {
            const long v = tid >= off ? s[tid - off] : 0;
            __syncthreads();
            s[tid] += v;
            __syncthreads();
        }
        if (b < nb) bsum[b] = carry + s[tid] - v0;
        __syncthreads();
        if (tid == FD_SCAN_BLOCK - 1) carry += s[tid];
        __syncthreads();
    }
    if (tid == 0) *total = carry;
}

__global__ __launch_bounds__(FD_SCAN_BLOCK) void fd_scan3_kernel(const long *local, const long *bsum,
                                                                 long nitems, const long *total,
                                                                 long *prefix)
{
    const long q = (long)blockIdx.x * FD_SCAN_BLOCK + threadIdx.x;
    if (q < nitems) prefix[q] = local[q] + bsum[blockIdx.x];
    if (q == 0) prefix[nitems] = *total;
}

#define FD_THREADS 256

// separate_modes: persistent walk over the flattened bins: workgroup b owns a
// contiguous range; each thread strides through it, reloading the item only
// when its next bin belongs to another (mode, segment).  Capped at 256 registers, two
// waves per SIMD: the per-bin chains of FP64 libm calls are latency-bound,
// and a second wave hides more of that than the few spills it costs
// (measured on full1-sized FD sums: 133 ms against 164 ms at one wave).
__global__ __launch_bounds__(FD_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2)))
void fd_sum_kernel(FdArgs a, long nitems)
{
    const long total = a.prefix[nitems];
    const long per = (total + gridDim.x - 1) / gridDim.x;
    const long w0 = (long)blockIdx.x * per;
    const long w1 = min(total, w0 + per);
    if (w0 >= w1) return;
    const fcx xi = fd_xi();
    const FdSpaConst kc = fd_spa_const();
    long item = -1;
    long item_end = 0, item_begin = 0;
    int lo = 0;
    FdItem it;
    for (long w = w0 + threadIdx.x; w < w1; w += FD_THREADS) {
        if (item < 0 || w >= item_end) {
            // last item with prefix[item] <= w (items without bins have equal prefixes)
            long l = item < 0 ? 0 : item, h = nitems;
            while (h - l > 1) {
                const long mid = (l + h) >> 1;
                if (a.prefix[mid] <= w) l = mid; else h = mid;
            }
            item = l;
            item_begin = a.prefix[item];
            item_end = a.prefix[item + 1];
            int hi;
            fd_item_range(a, item, lo, hi);
            fd_load_item(a, (int)(item / a.nseg), (int)(item % a.nseg), it);
        }
        const int i = lo + (int)(w - item_begin);
        const double f = a.freqs[i];
        const int q = (int)(item / a.nseg);
        double2 *out = a.wave + (long)q * a.data_length + i;
        // one writer per (mode, bin); the -f terms are not part of this output
        fd_bin(it, f, xi, kc, [&](fcx plus, fcx) { *out = make_double2(plus.re, plus.im); });
    }
}

// ------------------------------------------------ bin-owned accumulation
// Each output bin b is written by exactly one thread, which adds every
// contribution to it in a fixed order (no atomics, run-to-run identical).
// The +f and -f terms of one (item, bin) evaluation go to bins i and 2z - i
// (z = zero_index), so a thread owns a bin pair: owner o = b <= z with its
// mirror b' = 2z - b ("low" owners), or a bin b > 2z with no mirror partner
// ("high" owners).  A workgroup takes FD_OWN_TILE owners; for every mode it
// finds the segments whose bin range meets the tile's bins or their mirrors
// (each mode's segments sorted by first bin, with the running maximum of
// their last bins, so two binary searches bound the candidates), and every
// (item, bin) evaluation happens once, in the thread that owns the bin.

#define FD_OWN_TILE 256  // threads per workgroup
#define FD_OWN_BPT 8     // owners per thread

struct FdOwnIndex {
    int *cnt;   // (K,) segments with in-range bins
    int *slo;   // (K, nseg) their first bins, ascending (ties by segment)
    int *sj;    // (K, nseg) the segment of each sorted position
    int *pmx;   // (K, nseg) running maximum of the last bins in that order
};

__host__ __device__ inline size_t fd_align256(size_t x) { return (x + 255) & ~(size_t)255; }

// One thread per mode: insertion sort of its non-empty segments by first bin.
__global__ __launch_bounds__(256) void fd_mode_index_kernel(FdArgs a, FdOwnIndex ix)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= a.K) return;
    const long base = (long)q * a.nseg;
    int c = 0;
    for (int j = 0; j < a.nseg; j++) {
        int lo, hi;
        fd_item_range(a, base + j, lo, hi);
        if (hi < lo) continue;
        int p = c;
        while (p > 0 && ix.slo[base + p - 1] > lo) {
            ix.slo[base + p] = ix.slo[base + p - 1];
            ix.sj[base + p] = ix.sj[base + p - 1];
            p--;
        }
        ix.slo[base + p] = lo;
        ix.sj[base + p] = j;
        c++;
    }
    int mx = -1;
    for (int p = 0; p < c; p++) {
        int lo, hi;
        fd_item_range(a, base + ix.sj[base + p], lo, hi);
        mx = max(mx, hi);
        ix.pmx[base + p] = mx;
    }
    ix.cnt[q] = c;
}

// Sorted positions [s, e) of mode q that can meet bins [w0, w1]: first bin
// <= w1 and running maximum of last bins >= w0 (e <= s: none).
__device__ inline void fd_candidates(const FdOwnIndex &ix, long base, int c, int w0, int w1,
                                     int &s, int &e)
{
    s = 0;
    e = 0;
    if (c == 0 || w1 < w0 || ix.slo[base] > w1 || ix.pmx[base + c - 1] < w0) return;
    int l = 0, h = c;  // e: first position with slo > w1
    while (l < h) {
        const int mid = (l + h) >> 1;
        if (ix.slo[base + mid] <= w1) l = mid + 1; else h = mid;
    }
    e = l;
    l = 0;
    h = e;  // s: first position with pmx >= w0
    while (l < h) {
        const int mid = (l + h) >> 1;
        if (ix.pmx[base + mid] >= w0) h = mid; else l = mid + 1;
    }
    s = l;
}

__global__ __launch_bounds__(FD_OWN_TILE) __attribute__((amdgpu_waves_per_eu(2, 2)))
void fd_own_kernel(FdArgs a, FdOwnIndex ix, int n_low, int hb0, int tiles_low)
{
    constexpr int SPAN = FD_OWN_TILE * FD_OWN_BPT;  // owners per tile
    __shared__ int s_q[2 * FD_OWN_TILE], s_s[2 * FD_OWN_TILE], s_e[2 * FD_OWN_TILE];
    __shared__ int s_scan[FD_OWN_TILE];
    // the thread's accumulators: [0] its bins, [1] their mirrors
    __shared__ double2 s_acc[2][FD_OWN_BPT][FD_OWN_TILE];
    const int tid = threadIdx.x;
    const int z = a.zero_index, nf = a.data_length;
    // Low tiles in descending order: the frequencies nearest zero, where most
    // modes' bands overlap, carry the most work and start first.
    const bool low = (int)blockIdx.x < tiles_low;
    const int tile = low ? tiles_low - 1 - (int)blockIdx.x : (int)blockIdx.x;
    // this tile's bins [w0, w1] and their mirrors [m0, m1] (empty when m1 < m0)
    int w0, w1, m0 = 0, m1 = -1;
    if (low) {
        w0 = tile * SPAN;
        w1 = min(w0 + SPAN, n_low) - 1;
        m0 = max(2 * z - w1, 0);
        m1 = min(2 * z - w0, nf - 1);
    } else {
        w0 = hb0 + (tile - tiles_low) * SPAN;
        w1 = min(w0 + SPAN, nf) - 1;
    }
    // owner k of this thread: bin w0 + k * FD_OWN_TILE + tid and (low tiles) its mirror
#pragma unroll
    for (int k = 0; k < FD_OWN_BPT; k++) {
        s_acc[0][k][tid] = make_double2(0.0, 0.0);
        s_acc[1][k][tid] = make_double2(0.0, 0.0);
    }
    const fcx xi = fd_xi();
    const FdSpaConst kc = fd_spa_const();
    const bool minus_on = a.include_minus_m != 0;
    for (int q0 = 0; q0 < a.K; q0 += FD_OWN_TILE) {
        // candidate segment ranges of mode q0 + tid, compacted in mode order
        const int q = q0 + tid;
        int sa = 0, ea = 0, sb = 0, eb = 0;
        if (q < a.K) {
            const long base = (long)q * a.nseg;
            const int c = ix.cnt[q];
            fd_candidates(ix, base, c, w0, w1, sa, ea);
            fd_candidates(ix, base, c, m0, m1, sb, eb);
            if (ea > sa && eb > sb && sb <= ea && sa <= eb) {  // overlapping: one range
                sa = min(sa, sb);
                ea = max(ea, eb);
                eb = sb;
            }
        }
        const int na = (ea > sa ? 1 : 0) + (eb > sb ? 1 : 0);
        s_scan[tid] = na;
        __syncthreads();
        for (int off = 1; off < FD_OWN_TILE; off <<= 1) {
            const int v = tid >= off ? s_scan[tid - off] : 0;
            __syncthreads();
            s_scan[tid] += v;
            __syncthreads();
        }
        int pos = s_scan[tid] - na;
        const int nent = s_scan[FD_OWN_TILE - 1];
        if (ea > sa) {
            s_q[pos] = q;
            s_s[pos] = sa;
            s_e[pos] = ea;
            pos++;
        }
        if (eb > sb) {
            s_q[pos] = q;
            s_s[pos] = sb;
            s_e[pos] = eb;
        }
        __syncthreads();
        // The items of the list in order; each wave skips the items none of
        // its bins meet.  A thread evaluates up to 2 FD_OWN_BPT bins per item
        // (its bins and their mirrors), so one item load serves several
        // evaluations, all through one fd_bin body (a second inlined copy of
        // the large evaluation costs instruction-cache misses).
        for (int en = 0; en < nent; en++) {
            const int qe = __builtin_amdgcn_readfirstlane(s_q[en]);
            const int ps = __builtin_amdgcn_readfirstlane(s_s[en]);
            const int pe = __builtin_amdgcn_readfirstlane(s_e[en]);
            const long base = (long)qe * a.nseg;
            const bool conj_term = minus_on && a.m[qe] != 0;
            for (int p = ps; p < pe; p++) {
                const int j = __builtin_amdgcn_readfirstlane(ix.sj[base + p]);
                int lo, hi;
                fd_item_range(a, base + j, lo, hi);
                unsigned mask = 0;  // bit 2k: bin of owner k, bit 2k + 1: its mirror
#pragma unroll
                for (int k = 0; k < FD_OWN_BPT; k++) {
                    const int bk = w0 + k * FD_OWN_TILE + tid;
                    if (bk <= w1) {
                        if (bk >= lo && bk <= hi) mask |= 1u << (2 * k);
                        const int mk = 2 * z - bk;
                        if (low && mk != bk && mk < nf && mk >= lo && mk <= hi)
                            mask |= 2u << (2 * k);
                    }
                }
                if (!__any(mask != 0)) continue;
                FdItem it;
                fd_load_item(a, qe, j, it);
#pragma nounroll
                for (int slot = 0; slot < 2 * FD_OWN_BPT; slot++) {
                    const bool act = (mask >> slot) & 1u;
                    if (!__any(act)) continue;
                    if (act) {
                        const int k = slot >> 1;
                        const bool mirror = slot & 1;
                        const int bk = w0 + k * FD_OWN_TILE + tid;
                        const int mk = 2 * z - bk;
                        // mirror == 0: +f to bk, -f to 2z - bk; mirror == 1: +f to the
                        // mirror bin, -f back to bk
                        fd_bin(it, a.freqs[mirror ? mk : bk], xi, kc, [&](fcx plus, fcx minus) {
                            double2 &tp = s_acc[mirror ? 1 : 0][k][tid];
                            tp.x += plus.re;
                            tp.y += plus.im;
                            if (conj_term) {
                                if (mirror || mk == bk) {
                                    double2 &tm = s_acc[0][k][tid];
                                    tm.x += minus.re;
                                    tm.y += minus.im;
                                } else if (low && mk >= 0 && mk < nf) {
                                    double2 &tm = s_acc[1][k][tid];
                                    tm.x += minus.re;
                                    tm.y += minus.im;
                                }
                            }
                        });
                    }
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < FD_OWN_BPT; k++) {
        const int bk = w0 + k * FD_OWN_TILE + tid;
        if (bk > w1) continue;
        const double2 ab = s_acc[0][k][tid];
        const double2 w = a.wave[bk];
        a.wave[bk] = make_double2(w.x + ab.x, w.y + ab.y);
        const int mk = 2 * z - bk;
        if (low && mk != bk && mk < nf) {
            const double2 am = s_acc[1][k][tid];
            const double2 wm = a.wave[mk];
            a.wave[mk] = make_double2(wm.x + am.x, wm.y + am.y);
        }
    }
}

// Owners and tiles of the bin-owned sum.
inline void fd_owner_split(int z, int nf, int &n_low, int &hb0, int &tiles_low, int &tiles)
{
    n_low = z >= 0 ? min(z, nf - 1) + 1 : 0;
    hb0 = (int)std::max(2L * z + 1, 0L);
    const int n_high = hb0 < nf ? nf - hb0 : 0;
    const int span = FD_OWN_TILE * FD_OWN_BPT;
    tiles_low = (n_low + span - 1) / span;
    tiles = tiles_low + (n_high + span - 1) / span;
}

static size_t fd_scan_ws_bytes(int K, int nseg)
{
    const long nitems = (long)K * nseg;
    const long nb = (nitems + FD_SCAN_BLOCK - 1) / FD_SCAN_BLOCK;
    // local (nitems) + prefix (nitems + 1) + block sums (nb) + total
    return sizeof(long) * (size_t)(2 * nitems + 1 + nb + 1);
}

size_t fd_workspace_bytes(int K, int nseg)
{
    const size_t nitems = (size_t)K * nseg;
    // the item-walk scan (separate_modes) and the bin-owned index share nothing
    return fd_align256(fd_scan_ws_bytes(K, nseg)) + fd_align256(sizeof(int) * (size_t)K) +
           3 * fd_align256(sizeof(int) * nitems);
}

int launch_fd_sum(FdArgs a, void *workspace, size_t workspace_bytes, hipStream_t stream)
{
    const long nitems = (long)a.K * a.nseg;
    if (nitems <= 0) return FEW_OK;
    if (workspace_bytes < fd_workspace_bytes(a.K, a.nseg)) {
        set_error("fd_sum: workspace of %zu bytes, need %zu", workspace_bytes,
                  fd_workspace_bytes(a.K, a.nseg));
        return FEW_ERR_ARG;
    }
    const long nb = (nitems + FD_SCAN_BLOCK - 1) / FD_SCAN_BLOCK;
    if (nb > 0x7fffffffL) {
        set_error("fd_sum: too many (mode, segment) items (%ld)", nitems);
        return FEW_ERR_LIMIT;
    }
    if (!a.separate_modes) {
        // bin-owned: every bin written once, contributions in a fixed order
        char *w = static_cast<char *>(workspace) + fd_align256(fd_scan_ws_bytes(a.K, a.nseg));
        FdOwnIndex ix;
        ix.cnt = reinterpret_cast<int *>(w);
        w += fd_align256(sizeof(int) * (size_t)a.K);
        ix.slo = reinterpret_cast<int *>(w);
        w += fd_align256(sizeof(int) * (size_t)nitems);
        ix.sj = reinterpret_cast<int *>(w);
        w += fd_align256(sizeof(int) * (size_t)nitems);
        ix.pmx = reinterpret_cast<int *>(w);
        hipLaunchKernelGGL(fd_mode_index_kernel, dim3((a.K + 255) / 256), dim3(256), 0, stream, a,
                           ix);
        FEW_HIP(hipGetLastError());
        int n_low, hb0, tiles_low, tiles;
        fd_owner_split(a.zero_index, a.data_length, n_low, hb0, tiles_low, tiles);
        if (tiles > 0) {
            hipLaunchKernelGGL(fd_own_kernel, dim3(tiles), dim3(FD_OWN_TILE), 0, stream, a, ix,
                               n_low, hb0, tiles_low);
            FEW_HIP(hipGetLastError());
        }
        return FEW_OK;
    }
    long *local = static_cast<long *>(workspace);
    long *prefix = local + nitems;
    long *bsum = prefix + nitems + 1;
    long *total = bsum + nb;
    hipLaunchKernelGGL(fd_scan1_kernel, dim3((unsigned)nb), dim3(FD_SCAN_BLOCK), 0, stream, a,
                       nitems, local, bsum);
    FEW_HIP(hipGetLastError());
    hipLaunchKernelGGL(fd_scan2_kernel, dim3(1), dim3(FD_SCAN_BLOCK), 0, stream, bsum, nb, total);
    FEW_HIP(hipGetLastError());
    hipLaunchKernelGGL(fd_scan3_kernel, dim3((unsigned)nb), dim3(FD_SCAN_BLOCK), 0, stream, local,
                       bsum, nitems, total, prefix);
    FEW_HIP(hipGetLastError());
    a.prefix = prefix;
    int dev = 0, ncu = 256;
    FEW_HIP(hipGetDevice(&dev));
    FEW_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    // 64 workgroups per CU, each a contiguous bin range: short ranges balance
    // the uneven per-bin cost (roots in or out of the segment, the SPA branch)
    // better than a few long ones (full1-sized FD sum: 124 ms at 64 per CU,
    // 133 ms at 8, 138 ms at 2)
    hipLaunchKernelGGL(fd_sum_kernel, dim3((unsigned)(ncu * 64)), dim3(FD_THREADS), 0, stream, a,
                       nitems);
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

}  // namespace few

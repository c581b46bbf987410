/*
 * few_oracle_fd.c -- CPU FP64 restatement of FastEMRIWaveforms' frequency-
 * domain interpolated sum (the CPU branch of `get_waveform_generic_fd`,
 * src/few/cutils/interpolate.cu:1142-1563), stationary-phase approximation.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only the test suite, the smoke
 * check of __graft_entry__ and the cpu_baseline leg of bench.py may load
 * this library, and only as the checker / CPU baseline.  The shipped
 * library (fastemriwaveforms_amd/csrc) never links or calls it.
 *
 * Built with -ffp-contract=off: every expression rounds as written.
 *
 * Complex arithmetic: the reference's `cmplx` is gcmplx::complex<double>
 * (src/few/cutils/cuda_complex.hpp), the libc++ <complex> algorithms.  The
 * helpers below restate the ones the FD kernel reaches, with their C99
 * Annex G special-value handling, on glibc's libm.
 */

#define _GNU_SOURCE 1
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

/* gcmplx::exp and polar take cos and sin of one argument; an optimising
 * compiler merges the pair into one sincos call, whose last bit can differ
 * from separate sin and cos calls.  The restatement calls sincos explicitly
 * so the result does not depend on this file's compiler or flags. */

typedef struct {
    double re, im;
} ocx;

static inline ocx ocx_make(double re, double im)
{
    ocx z = {re, im};
    return z;
}

/* cuda_complex.hpp:440-498 -- product with NaN recovery */
static ocx ocx_mul(ocx z, ocx w)
{
    double a = z.re, b = z.im, c = w.re, d = w.im;
    double ac = a * c, bd = b * d, ad = a * d, bc = b * c;
    double x = ac - bd, y = ad + bc;
    if (isnan(x) && isnan(y)) {
        int recalc = 0;
        if (isinf(a) || isinf(b)) {
            a = copysign(isinf(a) ? 1.0 : 0.0, a);
            b = copysign(isinf(b) ? 1.0 : 0.0, b);
            if (isnan(c)) c = copysign(0.0, c);
            if (isnan(d)) d = copysign(0.0, d);
            recalc = 1;
        }
        if (isinf(c) || isinf(d)) {
            c = copysign(isinf(c) ? 1.0 : 0.0, c);
            d = copysign(isinf(d) ? 1.0 : 0.0, d);
            if (isnan(a)) a = copysign(0.0, a);
            if (isnan(b)) b = copysign(0.0, b);
            recalc = 1;
        }
        if (!recalc && (isinf(ac) || isinf(bd) || isinf(ad) || isinf(bc))) {
            if (isnan(a)) a = copysign(0.0, a);
            if (isnan(b)) b = copysign(0.0, b);
            if (isnan(c)) c = copysign(0.0, c);
            if (isnan(d)) d = copysign(0.0, d);
            recalc = 1;
        }
        if (recalc) {
            x = INFINITY * (a * c - b * d);
            y = INFINITY * (a * d + b * c);
        }
    }
    return ocx_make(x, y);
}

/* complex * real and complex / real: component-wise (cuda_complex.hpp:224-225,
 * 500-518, 618-621) */
static inline ocx ocx_scale(ocx z, double s) { return ocx_make(z.re * s, z.im * s); }
static inline ocx ocx_divr(ocx z, double s) { return ocx_make(z.re / s, z.im / s); }
static inline ocx ocx_add(ocx z, ocx w) { return ocx_make(z.re + w.re, z.im + w.im); }
/* real + complex (cuda_complex.hpp:404-409): the real part only */
static inline ocx ocx_addr(double r, ocx z) { return ocx_make(z.re + r, z.im); }

/* cuda_complex.hpp operator/(complex, complex): logb/scalbn-scaled quotient */
static ocx ocx_div(ocx z, ocx w)
{
    int ilogbw = 0;
    double a = z.re, b = z.im, c = w.re, d = w.im;
    double logbw = logb(fmax(fabs(c), fabs(d)));
    if (isfinite(logbw)) {
        ilogbw = (int)logbw;
        c = scalbn(c, -ilogbw);
        d = scalbn(d, -ilogbw);
    }
    double denom = c * c + d * d;
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
    return ocx_make(x, y);
}

/* real / complex (cuda_complex.hpp:623-630): complex(x) / w */
static inline ocx ocx_rdiv(double x, ocx w) { return ocx_div(ocx_make(x, 0.0), w); }

/* cuda_complex.hpp:750-765 -- abs = hypot, arg = atan2 */
static inline double ocx_abs(ocx z) { return hypot(z.re, z.im); }
static inline double ocx_arg(ocx z) { return atan2(z.im, z.re); }

/* cuda_complex.hpp:864-893 -- polar */
static ocx ocx_polar(double rho, double theta)
{
    if (isnan(rho) || signbit(rho)) return ocx_make(NAN, NAN);
    if (isnan(theta)) {
        if (isinf(rho)) return ocx_make(rho, theta);
        return ocx_make(theta, theta);
    }
    if (isinf(theta)) {
        if (isinf(rho)) return ocx_make(rho, NAN);
        return ocx_make(NAN, NAN);
    }
    double sn, cs;
    sincos(theta, &sn, &cs);
    double x = rho * cs;
    if (isnan(x)) x = 0;
    double y = rho * sn;
    if (isnan(y)) y = 0;
    return ocx_make(x, y);
}

/* cuda_complex.hpp:897-902 -- log */
static inline ocx ocx_log(ocx z) { return ocx_make(log(ocx_abs(z)), ocx_arg(z)); }

/* cuda_complex.hpp:914-930 -- sqrt */
static ocx ocx_sqrt(ocx z)
{
    if (isinf(z.im)) return ocx_make(INFINITY, z.im);
    if (isinf(z.re)) {
        if (z.re > 0.0) return ocx_make(z.re, isnan(z.im) ? z.im : copysign(0.0, z.im));
        return ocx_make(isnan(z.im) ? z.im : 0.0, copysign(z.re, z.im));
    }
    return ocx_polar(sqrt(ocx_abs(z)), ocx_arg(z) / 2.0);
}

/* cuda_complex.hpp:934-958 -- exp */
static ocx ocx_exp(ocx z)
{
    double i = z.im;
    if (isinf(z.re)) {
        if (z.re < 0.0) {
            if (!isfinite(i)) i = 1.0;
        } else if (i == 0 || !isfinite(i)) {
            if (isinf(i)) i = NAN;
            return ocx_make(z.re, i);
        }
    } else if (isnan(z.re) && z.im == 0) {
        return z;
    }
    double e = exp(z.re), sn, cs;
    sincos(i, &sn, &cs);
    return ocx_make(e * cs, e * sn);
}

/* cuda_complex.hpp:960-976 -- pow(complex, real) = exp(complex(y) * log(x)) */
static inline ocx ocx_pow(ocx x, double y) { return ocx_exp(ocx_mul(ocx_make(y, 0.0), ocx_log(x))); }

/* ------------------------------------------------------------------------ */
/* interpolate.cu:1041-1090 -- SPAFunc(x) = i sqrt(x) e^{-ix} K_{1/3}(-ix)   */
/* (series for |x| <= 7, asymptotic series beyond).                         */
/* ------------------------------------------------------------------------ */
static ocx orc_spa_func(double x)
{
    const ocx II = ocx_make(0.0, 1.0);
    const double Gamp13 = 2.67893853470774763;   /* Gamma(1/3) */
    const double Gamm13 = -4.06235381827920125;  /* Gamma(-1/3) */
    if (fabs(x) <= 7.) {
        const ocx xx = ocx_make(x, 0.0);
        const ocx tmp_p = ocx_pow(xx, 1. / 6.);
        /* exp(-2 pi i / 3) tmp_p^5 Gamma(-1/3) / 2^(1/3) */
        ocx e1 = ocx_exp(ocx_divr(ocx_scale(II, -2. * M_PI), 3.));
        ocx pref1 = ocx_mul(e1, tmp_p);
        pref1 = ocx_mul(pref1, tmp_p);
        pref1 = ocx_mul(pref1, tmp_p);
        pref1 = ocx_mul(pref1, tmp_p);
        pref1 = ocx_mul(pref1, tmp_p);
        pref1 = ocx_divr(ocx_scale(pref1, Gamm13), 1.2599210498948732);
        /* exp(-pi i / 3) tmp_p Gamma(1/3) / 2^(2/3) */
        ocx e2 = ocx_exp(ocx_divr(ocx_scale(II, -M_PI), 3.));
        ocx pref2 = ocx_mul(e2, tmp_p);
        pref2 = ocx_divr(ocx_scale(pref2, Gamp13), 1.5874010519681994748);
        const double x2 = x * x;
        const double c1_0 = 0.5, c1_2 = -0.09375, c1_4 = 0.0050223214285714285714;
        const double c1_6 = -0.00012555803571428571429, c1_8 = 1.8109332074175824176e-6;
        const double c1_10 = -1.6977498819539835165e-8, c1_12 = 1.1169407118118312608e-10;
        const double c1_14 = -5.4396463237589184781e-13, c1_16 = 2.0398673714095944293e-15;
        const double c1_18 = -6.0710338434809358015e-18, c1_20 = 1.4687985105195812423e-20;
        const double c1_22 = -2.9454515585285720100e-23, c1_24 = 4.9754249299469121790e-26;
        const double c1_26 = -7.1760936489618925658e-29;
        const double ser1 = c1_0 + x2 * (c1_2 + x2 * (c1_4 + x2 * (c1_6 + x2 * (c1_8 + x2 * (c1_10 + x2 * (c1_12 + x2 * (c1_14 + x2 * (c1_16 + x2 * (c1_18 + x2 * (c1_20 + x2 * (c1_22 + x2 * (c1_24 + x2 * c1_26))))))))))));
        const double c2_0 = 1., c2_2 = -0.375, c2_4 = 0.028125, c2_6 = -0.00087890625;
        const double c2_8 = 0.000014981356534090909091, c2_10 = -1.6051453429383116883e-7;
        const double c2_12 = 1.1802539286311115355e-9, c2_14 = -6.3227889033809546546e-12;
        const double c2_16 = 2.5772237377911499951e-14, c2_18 = -8.2603324929203525483e-17;
        const double c2_20 = 2.1362928861000911763e-19, c2_22 = -4.5517604107246260858e-22;
        const double c2_24 = 8.1281435905796894390e-25, c2_26 = -1.2340298973552160079e-27;
        const double ser2 = c2_0 + x2 * (c2_2 + x2 * (c2_4 + x2 * (c2_6 + x2 * (c2_8 + x2 * (c2_10 + x2 * (c2_12 + x2 * (c2_14 + x2 * (c2_16 + x2 * (c2_18 + x2 * (c2_20 + x2 * (c2_22 + x2 * (c2_24 + x2 * c2_26))))))))))));
        /* exp(-i x) (pref1 ser1 + pref2 ser2) */
        ocx ex = ocx_exp(ocx_scale(ocx_make(-II.re, -II.im), x));
        return ocx_mul(ex, ocx_add(ocx_scale(pref1, ser1), ocx_scale(pref2, ser2)));
    } else {
        const ocx y = ocx_make(1. / x, 0.0);
        ocx pref = ocx_exp(ocx_scale(ocx_scale(II, -0.75), M_PI));
        pref = ocx_scale(pref, sqrt(0.5 * M_PI));
        const ocx c_0 = II, c_1 = ocx_make(0.069444444444444444444, 0.0);
        const ocx c_2 = ocx_scale(II, -0.037133487654320987654);
        const ocx c_3 = ocx_make(-0.037993059127800640146, 0.0);
        const ocx c_4 = ocx_scale(II, 0.057649190412669721333);
        const ocx c_5 = ocx_make(0.11609906402551541102, 0.0);
        const ocx c_6 = ocx_scale(II, -0.29159139923075051147);
        const ocx c_7 = ocx_make(-0.87766696951001691647, 0.0);
        const ocx c_8 = ocx_scale(II, 3.0794530301731669934);
        /* c_0 + y (c_1 + y (c_2 + ... + y c_8)) */
        ocx ser = ocx_add(c_7, ocx_mul(y, c_8));
        ser = ocx_add(c_6, ocx_mul(y, ser));
        ser = ocx_add(c_5, ocx_mul(y, ser));
        ser = ocx_add(c_4, ocx_mul(y, ser));
        ser = ocx_add(c_3, ocx_mul(y, ser));
        ser = ocx_add(c_2, ocx_mul(y, ser));
        ser = ocx_add(c_1, ocx_mul(y, ser));
        ser = ocx_add(c_0, ocx_mul(y, ser));
        return ocx_mul(pref, ser);
    }
}

/* interpolate.cu:1113-1134 -- roots of a x^3 + b x^2 + c x + d (Cardano,
 * principal complex cube root).  The caller keeps only the real parts. */
static void orc_cube_roots(ocx *r1o, ocx *r2o, ocx *r3o, double a, double b, double c, double d)
{
    const double b2 = b * b;
    const double b3 = b2 * b;
    const double a2 = a * a;
    const double Delta_0 = (b2 - 3 * a * c);
    const double Delta_1 = (2 * b3 - 9 * a * b * c + 27 * a2 * d);
    const ocx sq = ocx_sqrt(ocx_make(Delta_1 * Delta_1 - 4 * (Delta_0 * Delta_0 * Delta_0), 0.0));
    const ocx C = ocx_pow(ocx_divr(ocx_addr(Delta_1, sq), 2.), 1. / 3.);
    const ocx xi = ocx_divr(ocx_addr(-1., ocx_sqrt(ocx_make(-3., 0.0))), 2.);
    const double pre = -(1. / (3 * a));
    const ocx xiC = ocx_mul(xi, C);
    const ocx xixiC = ocx_mul(ocx_mul(xi, xi), C);
    *r1o = ocx_scale(ocx_add(ocx_addr(b, C), ocx_rdiv(Delta_0, C)), pre);
    *r2o = ocx_scale(ocx_add(ocx_addr(b, xiC), ocx_rdiv(Delta_0, xiC)), pre);
    *r3o = ocx_scale(ocx_add(ocx_addr(b, xixiC), ocx_rdiv(Delta_0, xixiC)), pre);
}

/* The dense-output phase and its first two derivatives (interpolate.cu:1460-1474). */
static double orc_fd_phase(const double *r, double s, double s1)
{
    return r[0] + s * (r[1] + s1 * (r[2] + s * (r[3] + s1 * (r[4] + s * (r[5] + s1 * (r[6] + s * r[7]))))));
}

static double orc_fd_fdot(const double *r, double s, double s2, double s3, double s4, double s5,
                          double segwidth2)
{
    return (-2. * r[2] + r[3] * (2. - 6. * s) + r[4] * (2. - 12. * s + 12. * s2) +
            r[5] * (6. * s - 24. * s2 + 20. * s3) + r[6] * (6. * s - 36. * s2 + 60. * s3 - 30. * s4) +
            r[7] * (12. * s2 - 60. * s3 + 90. * s4 - 42. * s5)) /
           (2. * M_PI * segwidth2);
}

static double orc_fd_fddot(const double *r, double s, double s2, double s3, double s4,
                           double segwidth3)
{
    return (-6. * r[3] + r[4] * (-12. + 24. * s) + r[5] * (6. - 48. * s + 60. * s2) +
            r[6] * (6. - 72. * s + 180. * s2 - 120. * s3) +
            r[7] * (24. * s - 180. * s2 + 360. * s3 - 210. * s4)) /
           (2. * M_PI * segwidth3);
}

/* ------------------------------------------------------------------------ */
/* interpolate.cu:1142-1500 (CPU branch) -- make_generic_kerr_waveform_fd.   */
/*   waveform       (data_length,) complex, accumulated into (or, with       */
/*                  separate_modes, (num_modes, data_length) overwritten)    */
/*   interp_array   (4, 2K + 3, init_length): Re A, Im A, f_phi, f_theta,    */
/*                  f_r spline coefficients                                   */
/*   phase_coeffs   (8, 3, init_length - 1) dense-output phase coefficients   */
/*   seg_start/end  (num_modes, num_segments) inclusive frequency-index range */
/* Loop order: modes, segments, frequencies, roots -- the order the CPU      */
/* branch accumulates in.  Returns 0.                                         */
/* ------------------------------------------------------------------------ */
int orc_fd_waveform(double *waveform, const double *interp_array, const double *phase_interp_t,
                    const double *phase_coeffs, const int *m_arr, const int *k_arr,
                    const int *n_arr, int num_modes, const double *old_time_arr, int init_length,
                    int data_length, const double *frequencies, const int *seg_start_inds,
                    const int *seg_end_inds, int num_segments, const double *ylm_all,
                    int zero_index, int include_minus_m, int separate_modes)
{
    const long num_base = (2L * num_modes + 3) * init_length;
    const long num_phase_spline = (long)(init_length - 1) * 3;
    const ocx complexI = ocx_make(0.0, 1.0);
    for (int mode_i = 0; mode_i < num_modes; mode_i++) {
        const int m = m_arr[mode_i], k = k_arr[mode_i], n = n_arr[mode_i];
        const ocx ylm_p = ocx_make(ylm_all[2 * mode_i], ylm_all[2 * mode_i + 1]);
        const ocx ylm_m = ocx_make(ylm_all[2 * (num_modes + mode_i)],
                                   ylm_all[2 * (num_modes + mode_i) + 1]);
        for (int seg_i = 0; seg_i < num_segments; seg_i++) {
            const int start = seg_start_inds[mode_i * num_segments + seg_i];
            const int end = seg_end_inds[mode_i * num_segments + seg_i];
            const double t_start = old_time_arr[seg_i], t_end = old_time_arr[seg_i + 1];
            double ar[4], ai[4], fph[4], fth[4], fr[4], pp[8], pt[8], pr[8];
            for (int c = 0; c < 4; c++) {
                ar[c] = interp_array[c * num_base + (long)mode_i * init_length + seg_i];
                ai[c] = interp_array[c * num_base + (long)(num_modes + mode_i) * init_length + seg_i];
                fph[c] = interp_array[c * num_base + (long)(2 * num_modes) * init_length + seg_i];
                fth[c] = interp_array[c * num_base + (long)(2 * num_modes + 1) * init_length + seg_i];
                fr[c] = interp_array[c * num_base + (long)(2 * num_modes + 2) * init_length + seg_i];
            }
            for (int c = 0; c < 8; c++) {
                pp[c] = phase_coeffs[c * num_phase_spline + seg_i];
                pt[c] = phase_coeffs[c * num_phase_spline + init_length - 1 + seg_i];
                pr[c] = phase_coeffs[c * num_phase_spline + 2 * init_length - 2 + seg_i];
            }
            const double seg_width = phase_interp_t[seg_i + 1] - phase_interp_t[seg_i];
            const double f_y = m * fph[0] + k * fth[0] + n * fr[0];
            const double f_c1 = m * fph[1] + k * fth[1] + n * fr[1];
            const double f_c2 = m * fph[2] + k * fth[2] + n * fr[2];
            const double f_c3 = m * fph[3] + k * fth[3] + n * fr[3];
            for (int i = start; i <= end; i++) {
                if ((i > data_length - 1) || (i < 0)) continue;
                const double f = frequencies[i];
                const int diff = abs(zero_index - i);
                const int minus_idx = (i < zero_index) ? zero_index + diff : zero_index - diff;
                ocx root[3];
                orc_cube_roots(&root[0], &root[1], &root[2], f_c3, f_c2, f_c1, (f_y - f));
                for (int root_i = 0; root_i < 3; root_i++) {
                    const double root_here = root[root_i].re;
                    const double t = t_start + root_here;
                    if (!((t < t_end) && (t >= t_start))) continue;
                    const double x = root_here, x2 = x * x, x3 = x * x2;
                    const double R_re = ar[0] + ar[1] * x + ar[2] * x2 + ar[3] * x3;
                    const double R_im = ai[0] + ai[1] * x + ai[2] * x2 + ai[3] * x3;
                    const double s = (t - phase_interp_t[seg_i]) / seg_width;
                    const double s1 = 1.0 - s;
                    const double s2 = s * s, s3 = s * s2, s4 = s * s3, s5 = s * s4;
                    const double segwidth2 = seg_width * seg_width;
                    const double segwidth3 = seg_width * segwidth2;
                    const double Phi_phi = orc_fd_phase(pp, s, s1);
                    const double Phi_theta = orc_fd_phase(pt, s, s1);
                    const double Phi_r = orc_fd_phase(pr, s, s1);
                    const double fdot_phi = orc_fd_fdot(pp, s, s2, s3, s4, s5, segwidth2);
                    const double fdot_theta = orc_fd_fdot(pt, s, s2, s3, s4, s5, segwidth2);
                    const double fdot_r = orc_fd_fdot(pr, s, s2, s3, s4, s5, segwidth2);
                    const double fddot_phi = orc_fd_fddot(pp, s, s2, s3, s4, segwidth3);
                    const double fddot_theta = orc_fd_fddot(pt, s, s2, s3, s4, segwidth3);
                    const double fddot_r = orc_fd_fddot(pr, s, s2, s3, s4, segwidth3);
                    const double phase_term = m * Phi_phi + k * Phi_theta + n * Phi_r;
                    const double fdot = m * fdot_phi + k * fdot_theta + n * fdot_r;
                    const double fddot = m * fddot_phi + k * fddot_theta + n * fddot_r;
                    const ocx R_amp = ocx_make(R_re, R_im);
                    const double temp_arg = 2. * M_PI * fdot * fdot * fdot / (3. * fddot * fddot);
                    const double pre = -1.0 * fdot / fabs(fddot) * 2. / sqrt(3.);
                    const ocx amp_term2 = ocx_div(ocx_scale(orc_spa_func(temp_arg), pre),
                                                  ocx_sqrt(ocx_make(temp_arg, 0.0)));
                    const ocx ph = ocx_exp(ocx_scale(complexI, 2. * M_PI * f * t - phase_term));
                    const ocx temp_exp = ocx_mul(ocx_mul(R_amp, amp_term2), ph);
                    const ocx R_tmp = ocx_mul(ylm_p, temp_exp);
                    ocx L_tmp = ocx_make(0.0, 0.0);
                    const int minus = (m != 0.0) && include_minus_m;
                    if (minus) L_tmp = ocx_mul(ylm_m, ocx_make(temp_exp.re, -temp_exp.im));
                    if (!separate_modes) {
                        waveform[2L * i] += R_tmp.re;
                        waveform[2L * i + 1] += R_tmp.im;
                        if (minus && minus_idx >= 0 && minus_idx < data_length) {
                            waveform[2L * minus_idx] += L_tmp.re;
                            waveform[2L * minus_idx + 1] += L_tmp.im;
                        }
                    } else {
                        waveform[2L * ((long)mode_i * data_length + i)] = R_tmp.re;
                        waveform[2L * ((long)mode_i * data_length + i) + 1] = R_tmp.im;
                    }
                }
            }
        }
    }
    return 0;
}

/* Building blocks, exported for the pinning tests. */
void orc_fd_spa_func(const double *x, long n, double *out)
{
    for (long i = 0; i < n; i++) {
        const ocx r = orc_spa_func(x[i]);
        out[2 * i] = r.re;
        out[2 * i + 1] = r.im;
    }
}

void orc_fd_cube_roots(const double *abcd, long n, double *out)
{
    for (long i = 0; i < n; i++) {
        ocx r[3];
        orc_cube_roots(&r[0], &r[1], &r[2], abcd[4 * i], abcd[4 * i + 1], abcd[4 * i + 2],
                       abcd[4 * i + 3]);
        for (int j = 0; j < 3; j++) {
            out[6 * i + 2 * j] = r[j].re;
            out[6 * i + 2 * j + 1] = r[j].im;
        }
    }
}

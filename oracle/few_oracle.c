/*
 * few_oracle.c -- CPU FP64 restatement of FastEMRIWaveforms' interpolated
 * mode summation (the `interpolate_arrays` spline build and the CPU branch
 * of `get_waveform`).
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only the test suite, the smoke
 * check of __graft_entry__ and the cpu_baseline leg of bench.py may load
 * this library, and only as the checker / CPU baseline.  The shipped
 * library (fastemriwaveforms_amd/csrc) never links or calls it.
 *
 * Floating-point policy: built with -ffp-contract=off, so every expression
 * below rounds exactly as written (no fused multiply-add).  The FP64 HIP
 * parity kernels follow the same operation order for everything that feeds
 * the mode phase m*Phi_phi + k*Phi_theta + n*Phi_r, which is what makes a
 * rel-L2 of 1e-12 reachable at phases of 1e6-1e7 rad.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the FastEMRIWaveforms repository root).
 */

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* src/few/cutils/interpolate.cu:22 -- CPU chunk size of the mode loop.  The
 * CPU branch adds the partial sum of each 5000-mode chunk into the output. */
#define ORC_MODES_CHUNK 5000

/* ------------------------------------------------------------------------ */
/* LAPACK dgtsv (Reference-LAPACK 3.12, NRHS = 1), restated.                 */
/* The reference CPU backend solves each not-a-knot system with              */
/* LAPACKE_dgtsv (src/few/cutils/interpolate.cu:231-236).  Gaussian          */
/* elimination with partial pivoting (row interchange when |d_i| < |dl_i|),  */
/* then back substitution through the second super-diagonal fill-in.         */
/*   dl[0..n-2] : sub-diagonal  A[i+1][i]   (overwritten by fill-in)         */
/*   d [0..n-1] : diagonal                                                   */
/*   du[0..n-2] : super-diagonal A[i][i+1]                                   */
/*   b [0..n-1] : right-hand side, overwritten by the solution               */
/* Returns LAPACK's INFO (0, or i>0 when U(i,i) is exactly zero; then the    */
/* back substitution is skipped exactly as LAPACK does).                     */
/* ------------------------------------------------------------------------ */
int orc_dgtsv(int n, double *dl, double *d, double *du, double *b)
{
    if (n <= 0) return 0;
    int i;
    double fact, temp;
    for (i = 0; i < n - 2; i++) {
        if (fabs(d[i]) >= fabs(dl[i])) {
            if (d[i] != 0.0) {
                fact = dl[i] / d[i];
                d[i + 1] = d[i + 1] - fact * du[i];
                b[i + 1] = b[i + 1] - fact * b[i];
            } else {
                return i + 1;
            }
            dl[i] = 0.0;
        } else {
            fact = d[i] / dl[i];
            d[i] = dl[i];
            temp = d[i + 1];
            d[i + 1] = du[i] - fact * temp;
            dl[i] = du[i + 1];
            du[i + 1] = -fact * dl[i];
            du[i] = temp;
            temp = b[i];
            b[i] = b[i + 1];
            b[i + 1] = temp - fact * b[i + 1];
        }
    }
    if (n > 1) {
        i = n - 2;
        if (fabs(d[i]) >= fabs(dl[i])) {
            if (d[i] != 0.0) {
                fact = dl[i] / d[i];
                d[i + 1] = d[i + 1] - fact * du[i];
                b[i + 1] = b[i + 1] - fact * b[i];
            } else {
                return i + 1;
            }
        } else {
            fact = d[i] / dl[i];
            d[i] = dl[i];
            temp = d[i + 1];
            d[i + 1] = du[i] - fact * temp;
            du[i] = temp;
            temp = b[i];
            b[i] = b[i + 1];
            b[i + 1] = temp - fact * b[i + 1];
        }
    }
    if (d[n - 1] == 0.0) return n;

    b[n - 1] = b[n - 1] / d[n - 1];
    if (n > 1) b[n - 2] = (b[n - 2] - du[n - 2] * b[n - 1]) / d[n - 2];
    for (i = n - 3; i >= 0; i--)
        b[i] = (b[i] - du[i] * b[i + 1] - dl[i] * b[i + 2]) / d[i];
    return 0;
}

/* src/few/cutils/interpolate.cu:49-102 (prep_splines): one row of the
 * scipy-style not-a-knot system for knot i of one array. */
static void orc_prep_row(int i, int length, double *b, double *ud, double *diag,
                         double *ld, const double *x, const double *y)
{
    double dx1, dx2, d, slope1, slope2;
    if (i == length - 1) {
        dx1 = x[length - 2] - x[length - 3];
        dx2 = x[length - 1] - x[length - 2];
        d = x[length - 1] - x[length - 3];
        slope1 = (y[length - 2] - y[length - 3]) / dx1;
        slope2 = (y[length - 1] - y[length - 2]) / dx2;
        b[length - 1] = ((dx2 * dx2 * slope1 + (2 * d + dx2) * dx1 * slope2) / d);
        diag[length - 1] = dx1;
        ld[length - 1] = d;
        ud[length - 1] = 0.0;
    } else if (i == 0) {
        dx1 = x[1] - x[0];
        dx2 = x[2] - x[1];
        d = x[2] - x[0];
        slope1 = (y[1] - y[0]) / dx1;
        slope2 = (y[2] - y[1]) / dx2;
        b[0] = ((dx1 + 2 * d) * dx2 * slope1 + dx1 * dx1 * slope2) / d;
        diag[0] = dx2;
        ud[0] = d;
        ld[0] = 0.0;
    } else {
        dx1 = x[i] - x[i - 1];
        dx2 = x[i + 1] - x[i];
        slope1 = (y[i] - y[i - 1]) / dx1;
        slope2 = (y[i + 1] - y[i]) / dx2;
        b[i] = 3.0 * (dx2 * slope1 + dx1 * slope2);
        diag[i] = 2 * (dx1 + dx2);
        ud[i] = dx1;
        ld[i] = dx2;
    }
}

/* src/few/cutils/interpolate.cu:28-43 (fill_coefficients) */
static void orc_fill_coefficients(int i, const double *dydx, double dx, const double *y,
                                  double *c1, double *c2, double *c3)
{
    double slope = (y[i + 1] - y[i]) / dx;
    double dydx_i = dydx[i];
    double t = (dydx_i + dydx[i + 1] - 2 * slope) / dx;
    c1[i] = dydx_i;
    c2[i] = (slope - dydx_i) / dx - t;
    c3[i] = t / dx;
}

/*
 * src/few/cutils/interpolate.cu:284-325 (interpolate_arrays, CPU branch):
 * fill_B (:107-145) -> fit_wrap/dgtsv (:201-239) -> set_spline_constants
 * (:150-198) -> fill_final_derivs (:242-280).
 *   t_arr        : (ninterps, length)  knot times, one row per array
 *   interp_array : (4, ninterps, length); slot 0 holds y on entry,
 *                  slots 1..3 receive c1..c3
 *   B, ud, diag, ld : (ninterps, length) scratch
 * Returns the first non-zero dgtsv INFO (the reference ignores it).
 */
int orc_interpolate_arrays(const double *t_arr, double *interp_array, int ninterps,
                           int length, double *B, double *ud, double *diag, double *ld)
{
    int first_info = 0;
    if (ninterps <= 0 || length < 3) return -1;
    for (int s = 0; s < ninterps; s++) {
        size_t lead = (size_t)s * length;
        for (int i = 0; i < length; i++)
            orc_prep_row(i, length, &B[lead], &ud[lead], &diag[lead], &ld[lead],
                         &t_arr[lead], &interp_array[lead]);
    }
    for (int s = 0; s < ninterps; s++) {
        size_t lead = (size_t)s * length;
        int info = orc_dgtsv(length, &ld[lead + 1], &diag[lead], &ud[lead], &B[lead]);
        if (info != 0 && first_info == 0) first_info = info;
    }
    size_t nb = (size_t)ninterps * length;
    for (int s = 0; s < ninterps; s++) {
        size_t lead = (size_t)s * length;
        for (int i = 0; i < length - 1; i++) {
            double dt = t_arr[lead + i + 1] - t_arr[lead + i];
            orc_fill_coefficients(i, &B[lead], dt, &interp_array[0 * nb + lead],
                                  &interp_array[1 * nb + lead], &interp_array[2 * nb + lead],
                                  &interp_array[3 * nb + lead]);
        }
    }
    for (int s = 0; s < ninterps; s++) {
        size_t lead = (size_t)s * length;
        double c1 = interp_array[1 * nb + lead + length - 2];
        double c2 = interp_array[2 * nb + lead + length - 2];
        double c3 = interp_array[3 * nb + lead + length - 2];
        double x = t_arr[lead + length - 1] - t_arr[lead + length - 2];
        double x2 = x * x;
        interp_array[1 * nb + lead + length - 1] = c1 + 2 * c2 * x + 3 * c3 * x2;
        interp_array[2 * nb + lead + length - 1] = (2. * c2 + 6. * c3 * x) / 2.;
        interp_array[3 * nb + lead + length - 1] = c3;
    }
    return first_info;
}

/*
 * src/few/cutils/interpolate.cu:694-723 (find_start_inds).  start_inds[i] is
 * the first output sample of knot interval i; a sample exactly on a knot
 * belongs to the later interval.  Returns the number of knots actually used
 * (the reference's updated *length).  When no knot reaches (new_length-1)*dt
 * the reference would go on to read start_inds[length] (never written); here
 * the count is capped at `length`, so samples past the last knot stay
 * untouched.
 */
int orc_find_start_inds(int *start_inds, const double *t_arr, double delta_t, int length,
                        int new_length)
{
    double T = (new_length - 1) * delta_t;
    start_inds[0] = 0;
    int i;
    for (i = 1; i < length; i++) {
        double t = t_arr[i];
        if (t < T) {
            start_inds[i] = (int)ceil(t / delta_t);
        } else {
            start_inds[i] = new_length;
            break;
        }
    }
    return (i + 1 > length) ? length : i + 1;
}

/* Complex product with the operation order of gcmplx::operator*
 * (src/few/cutils/cuda_complex.hpp:444-466); the C99 NaN/Inf recovery
 * branch never triggers on finite input. */
static inline void cmul(double a, double b, double c, double d, double *re, double *im)
{
    double ac = a * c, bd = b * d, ad = a * d, bc = b * c;
    *re = ac - bd;
    *im = ad + bc;
}

/*
 * Phase of one DOPR853 dense-output segment: src/few/cutils/interpolate.cu:642-644
 * (same nested form as src/few/trajectory/dopr853.py:795-803).
 * coeffs are the 8 rcont values of one (interval, phase) pair.
 */
static inline double orc_dense_phase(const double *r, double s, double s1)
{
    return r[0] + s * (r[1] + s1 * (r[2] + s * (r[3] + s1 * (r[4] + s * (r[5] + s1 * (r[6] + s * r[7]))))));
}

typedef struct {
    const double *interp_array;   /* (4, init_len, 2K) */
    const double *phase_t;        /* (init_len,) */
    const double *phase_coeffs;   /* (8, init_len-1, 3) */
    const int *m, *k, *n;
    int init_len, num_modes;
    const double *ylms;           /* 2K complex: Y+ then Y- */
    double delta_t;
    const double *h_t;            /* (init_len,) */
} orc_wave_args;

/*
 * One output sample i of knot interval j: the body of make_waveform's CPU
 * branch (src/few/cutils/interpolate.cu:618-686).  Writes the value the
 * reference would leave in a zero-initialised waveform[i].
 */
static void orc_eval_sample(const orc_wave_args *a, long i, int j, double *out_re, double *out_im)
{
    const int K = a->num_modes;
    const long num_base = (long)a->init_len * (2 * K);
    const long num_phase = (long)(a->init_len - 1) * 3;
    double r_phi[8], r_theta[8], r_r[8];
    for (int c = 0; c < 8; c++) {
        r_phi[c] = a->phase_coeffs[c * num_phase + j * 3 + 0];
        r_theta[c] = a->phase_coeffs[c * num_phase + j * 3 + 1];
        r_r[c] = a->phase_coeffs[c * num_phase + j * 3 + 2];
    }
    const double start_t = a->h_t[j];
    const double pt_here = a->phase_t[j];
    const double seg = a->phase_t[j + 1] - a->phase_t[j];

    double t = a->delta_t * i;
    double x = t - start_t;
    double x2 = x * x;
    double x3 = x * x2;
    double s = (t - pt_here) / seg;
    double s1 = 1.0 - s;
    double Phi_phi = orc_dense_phase(r_phi, s, s1);
    double Phi_theta = orc_dense_phase(r_theta, s, s1);
    double Phi_r = orc_dense_phase(r_r, s, s1);

    double acc_re = 0.0, acc_im = 0.0;
    for (int c0 = 0; c0 < K; c0 += ORC_MODES_CHUNK) {
        int c1 = c0 + ORC_MODES_CHUNK < K ? c0 + ORC_MODES_CHUNK : K;
        double tr_re = 0.0, tr_im = 0.0;
        for (int mode = c0; mode < c1; mode++) {
            const long ind_re = (long)j * (2 * K) + mode;
            const long ind_im = ind_re + K;
            double re_val = a->interp_array[ind_re] + a->interp_array[num_base + ind_re] * x +
                            a->interp_array[2 * num_base + ind_re] * x2 +
                            a->interp_array[3 * num_base + ind_re] * x3;
            double im_val = a->interp_array[ind_im] + a->interp_array[num_base + ind_im] * x +
                            a->interp_array[2 * num_base + ind_im] * x2 +
                            a->interp_array[3 * num_base + ind_im] * x3;
            const int m = a->m[mode], kk = a->k[mode], n = a->n[mode];
            double phase = (double)m * Phi_phi + (double)kk * Phi_theta + (double)n * Phi_r;
            /* exp(minus_I * phase) = (cos(-phase), sin(-phase))
             * (cuda_complex.hpp:937-958 with a zero real part) */
            double e_re = cos(-phase), e_im = sin(-phase);
            double p_re, p_im;
            cmul(e_re, e_im, re_val, im_val, &p_re, &p_im);
            const double yp_re = a->ylms[2 * mode], yp_im = a->ylms[2 * mode + 1];
            double tp_re, tp_im;
            cmul(p_re, p_im, yp_re, yp_im, &tp_re, &tp_im);
            double tm_re = 0.0, tm_im = 0.0;
            if (m != 0) {
                const double ym_re = a->ylms[2 * (K + mode)], ym_im = a->ylms[2 * (K + mode) + 1];
                cmul(p_re, -p_im, ym_re, ym_im, &tm_re, &tm_im);
            }
            tr_re = tr_re + tm_re + tp_re;
            tr_im = tr_im + tm_im + tp_im;
        }
        acc_re += tr_re;
        acc_im += tr_im;
    }
    *out_re = acc_re;
    *out_im = acc_im;
}

static void orc_fill_args(orc_wave_args *a, const double *interp_array, const double *phase_t,
                          const double *phase_coeffs, const int *m, const int *k, const int *n,
                          int init_len, int num_modes, const double *ylms, double delta_t,
                          const double *h_t)
{
    a->interp_array = interp_array;
    a->phase_t = phase_t;
    a->phase_coeffs = phase_coeffs;
    a->m = m;
    a->k = k;
    a->n = n;
    a->init_len = init_len;
    a->num_modes = num_modes;
    a->ylms = ylms;
    a->delta_t = delta_t;
    a->h_t = h_t;
}

/*
 * src/few/cutils/interpolate.cu:726-797 (get_waveform, CPU branch).
 * waveform (out_len complex, interleaved re/im) is accumulated into.
 * nthreads <= 1 runs serially (the reference CPU backend is single-threaded);
 * larger values split samples across OpenMP threads.  Every sample is
 * computed by the same code either way, so results are bit-identical.
 */
int orc_get_waveform(double *waveform, const double *interp_array, const double *phase_t,
                     const double *phase_coeffs, const int *m, const int *k, const int *n,
                     int init_len, int out_len, int num_modes, const double *ylms,
                     double delta_t, const double *h_t, int nthreads)
{
    if (init_len < 2 || out_len < 0 || num_modes < 0) return -1;
    int *start_inds = (int *)malloc(sizeof(int) * (size_t)(init_len + 1));
    if (!start_inds) return -2;
    int used = orc_find_start_inds(start_inds, h_t, delta_t, init_len, out_len);
    orc_wave_args a;
    orc_fill_args(&a, interp_array, phase_t, phase_coeffs, m, k, n, init_len, num_modes, ylms,
                  delta_t, h_t);
    for (int j = 0; j < used - 1; j++) {
        long lo = start_inds[j], hi = start_inds[j + 1];
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
#endif
        for (long i = lo; i < hi; i++) {
            double re, im;
            orc_eval_sample(&a, i, j, &re, &im);
            waveform[2 * i] += re;
            waveform[2 * i + 1] += im;
        }
    }
    free(start_inds);
    return 0;
}

/*
 * The same per-sample arithmetic at an arbitrary list of sample indices:
 * out[2q], out[2q+1] = the value get_waveform would leave in a zeroed
 * waveform[idx[q]] (or 0 when the sample lies in no interval).  Lets tests
 * check full-size outputs at a few thousand samples in seconds.
 */
int orc_waveform_samples(double *out, const int64_t *idx, long nidx, const double *interp_array,
                         const double *phase_t, const double *phase_coeffs, const int *m,
                         const int *k, const int *n, int init_len, int out_len, int num_modes,
                         const double *ylms, double delta_t, const double *h_t, int nthreads)
{
    if (init_len < 2 || out_len < 0 || num_modes < 0) return -1;
    int *start_inds = (int *)malloc(sizeof(int) * (size_t)(init_len + 1));
    if (!start_inds) return -2;
    int used = orc_find_start_inds(start_inds, h_t, delta_t, init_len, out_len);
    orc_wave_args a;
    orc_fill_args(&a, interp_array, phase_t, phase_coeffs, m, k, n, init_len, num_modes, ylms,
                  delta_t, h_t);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
#endif
    for (long q = 0; q < nidx; q++) {
        long i = (long)idx[q];
        int j = -1;
        for (int jj = 0; jj < used - 1; jj++)
            if (i >= start_inds[jj] && i < start_inds[jj + 1]) { j = jj; break; }
        double re = 0.0, im = 0.0;
        if (j >= 0) orc_eval_sample(&a, i, j, &re, &im);
        out[2 * q] = re;
        out[2 * q + 1] = im;
    }
    free(start_inds);
    return 0;
}

/*
 * Conversion of one cubic piece y + c1 x + c2 x^2 + c3 x^3 on an interval of
 * width h into the 8-coefficient dense-output form evaluated at
 * interpolate.cu:642 (r5..r8 = 0).  With s = x/h and s1 = 1-s the nested form
 * r1 + s(r2 + s1(r3 + s r4)) expands to r1 + s(r2+r3) + s^2(r4-r3) - s^3 r4,
 * which equals the cubic when r2 = c1 h + c2 h^2 + c3 h^3,
 * r3 = -(c2 h^2 + c3 h^3) and r4 = -c3 h^3.  This is the format adapter for
 * synthetic phase knots (no reference counterpart; the reference gets these
 * coefficients from its ODE integrator, trajectory/dopr853.py:621-758).
 *   coeffs : (4, nphase, length) cubic coefficients as interpolate_arrays
 *            leaves them, t : (length,) knots shared by the phase arrays
 *   out    : (8, length-1, nphase)  -- the layout get_waveform reads
 */
void orc_cubic_to_dense(const double *coeffs, const double *t, int nphase, int length, double *out)
{
    const long nb = (long)nphase * length;
    const long no = (long)(length - 1) * nphase;
    for (int j = 0; j < length - 1; j++) {
        double h = t[j + 1] - t[j];
        double h2 = h * h;
        double h3 = h2 * h;
        for (int p = 0; p < nphase; p++) {
            long src = (long)p * length + j;
            double y = coeffs[src];
            double q1 = coeffs[nb + src] * h;
            double q2 = coeffs[2 * nb + src] * h2;
            double q3 = coeffs[3 * nb + src] * h3;
            long dst = (long)j * nphase + p;
            out[0 * no + dst] = y;
            out[1 * no + dst] = (q1 + q2) + q3;
            out[2 * no + dst] = -(q2 + q3);
            out[3 * no + dst] = -q3;
            for (int c = 4; c < 8; c++) out[c * no + dst] = 0.0;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Adjoint of the mode sum and of the spline build.                          */
/* The reference has no backward pass; these restate the exact derivatives  */
/* of orc_get_waveform and orc_interpolate_arrays above, for the loss        */
/*   L = sum_i Re(conj(g_i) h_i) = sum_i (g_re h_re + g_im h_im),            */
/* g being the upstream gradient dL/dh (interleaved complex).                */
/* ------------------------------------------------------------------------ */

/*
 * Gradient of L with respect to the coefficients the sum reads
 * (make_waveform, src/few/cutils/interpolate.cu:618-686, restated in
 * orc_eval_sample).  For sample i of interval j with x = t - h_t[j],
 * e = exp(-i phi), P = e A, T = P Y+ + [m != 0] conj(P) Y-:
 *   dL/dRe A = Re(q+ Y+) + [m != 0] Re(q- Y-)
 *   dL/dIm A = -Im(q+ Y+) + [m != 0] Im(q- Y-)
 *   dL/dphi  = Im(q+ A Y+) - [m != 0] Im(q- conj(A) Y-)
 * with q+ = conj(g) e and q- = conj(g) conj(e); phi = m Phi_phi + k Phi_theta
 * + n Phi_r, so dL/dPhi_phi = sum_modes m dL/dphi (k, n likewise).  The
 * phase of interval j is the cubic sum_p c_p xp^p, xp = t - phase_t[j]
 * (the dense-output form evaluated at :642 equals it identically, see
 * orc_cubic_to_dense), so dL/dc_p = sum_i xp^p dL/dPhi.
 *   g_interp : (4, init_len, 2K), the interp_array layout (slot p, knot j,
 *              Re part at mode, Im part at K + mode); row init_len-1 and
 *              intervals without samples stay zero
 *   g_phase  : (4, init_len, 3): dL/dc_p of phase q on interval j at
 *              [p][j][q] (cubic coefficients of the phase splines)
 * Both outputs are overwritten.  Intervals are independent, so OpenMP
 * splits them; each interval sums its samples in order.
 */
int orc_waveform_adjoint(const double *grad_wave, const double *interp_array,
                         const double *phase_t, const double *phase_coeffs, const int *m,
                         const int *k, const int *n, int init_len, int out_len, int num_modes,
                         const double *ylms, double delta_t, const double *h_t,
                         double *g_interp, double *g_phase, int nthreads)
{
    if (init_len < 2 || out_len < 0 || num_modes < 0) return -1;
    const int K = num_modes;
    const long num_base = (long)init_len * (2 * K);
    const long num_phase = (long)(init_len - 1) * 3;
    memset(g_interp, 0, sizeof(double) * 4 * (size_t)num_base);
    memset(g_phase, 0, sizeof(double) * 4 * (size_t)init_len * 3);
    int *start_inds = (int *)malloc(sizeof(int) * (size_t)(init_len + 1));
    if (!start_inds) return -2;
    int used = orc_find_start_inds(start_inds, h_t, delta_t, init_len, out_len);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
#endif
    for (int j = 0; j < used - 1; j++) {
        double r_phi[8], r_theta[8], r_r[8];
        for (int c = 0; c < 8; c++) {
            r_phi[c] = phase_coeffs[c * num_phase + j * 3 + 0];
            r_theta[c] = phase_coeffs[c * num_phase + j * 3 + 1];
            r_r[c] = phase_coeffs[c * num_phase + j * 3 + 2];
        }
        const double start_t = h_t[j];
        const double pt_here = phase_t[j];
        const double seg = phase_t[j + 1] - phase_t[j];
        double gp[4][3] = {{0}};
        for (long i = start_inds[j]; i < start_inds[j + 1]; i++) {
            const double t = delta_t * i;
            const double x = t - start_t;
            const double xpow[4] = {1.0, x, x * x, x * x * x};
            const double s = (t - pt_here) / seg;
            const double s1 = 1.0 - s;
            const double Phi_phi = orc_dense_phase(r_phi, s, s1);
            const double Phi_theta = orc_dense_phase(r_theta, s, s1);
            const double Phi_r = orc_dense_phase(r_r, s, s1);
            const double gr = grad_wave[2 * i], gi = grad_wave[2 * i + 1];
            double dphi = 0.0, dtheta = 0.0, dr = 0.0;
            for (int mode = 0; mode < K; mode++) {
                const long ind_re = (long)j * (2 * K) + mode;
                const long ind_im = ind_re + K;
                double a_re = interp_array[ind_re] + interp_array[num_base + ind_re] * x +
                              interp_array[2 * num_base + ind_re] * xpow[2] +
                              interp_array[3 * num_base + ind_re] * xpow[3];
                double a_im = interp_array[ind_im] + interp_array[num_base + ind_im] * x +
                              interp_array[2 * num_base + ind_im] * xpow[2] +
                              interp_array[3 * num_base + ind_im] * xpow[3];
                const int mm = m[mode], kk = k[mode], nn = n[mode];
                double phase = (double)mm * Phi_phi + (double)kk * Phi_theta + (double)nn * Phi_r;
                double e_re = cos(-phase), e_im = sin(-phase);
                double qp_re, qp_im, qm_re, qm_im;
                cmul(gr, -gi, e_re, e_im, &qp_re, &qp_im);
                cmul(gr, -gi, e_re, -e_im, &qm_re, &qm_im);
                const double yp_re = ylms[2 * mode], yp_im = ylms[2 * mode + 1];
                double u_re, u_im;
                cmul(qp_re, qp_im, yp_re, yp_im, &u_re, &u_im);       /* q+ Y+ */
                double d_re = u_re, d_im = -u_im;
                double ay_re, ay_im, w_re, w_im;
                cmul(a_re, a_im, yp_re, yp_im, &ay_re, &ay_im);       /* A Y+ */
                cmul(qp_re, qp_im, ay_re, ay_im, &w_re, &w_im);
                double rho = w_im;
                if (mm != 0) {
                    const double ym_re = ylms[2 * (K + mode)], ym_im = ylms[2 * (K + mode) + 1];
                    double v_re, v_im;
                    cmul(qm_re, qm_im, ym_re, ym_im, &v_re, &v_im);   /* q- Y- */
                    d_re += v_re;
                    d_im += v_im;
                    cmul(a_re, -a_im, ym_re, ym_im, &ay_re, &ay_im);  /* conj(A) Y- */
                    cmul(qm_re, qm_im, ay_re, ay_im, &w_re, &w_im);
                    rho -= w_im;
                }
                for (int p = 0; p < 4; p++) {
                    g_interp[p * num_base + ind_re] += xpow[p] * d_re;
                    g_interp[p * num_base + ind_im] += xpow[p] * d_im;
                }
                dphi += (double)mm * rho;
                dtheta += (double)kk * rho;
                dr += (double)nn * rho;
            }
            const double xp = t - pt_here;
            double xpp = 1.0;
            for (int p = 0; p < 4; p++) {
                gp[p][0] += xpp * dphi;
                gp[p][1] += xpp * dtheta;
                gp[p][2] += xpp * dr;
                xpp *= xp;
            }
        }
        for (int p = 0; p < 4; p++)
            for (int q = 0; q < 3; q++) g_phase[((long)p * init_len + j) * 3 + q] = gp[p][q];
    }
    free(start_inds);
    return 0;
}

/*
 * Adjoint (transpose) of the linear map y -> interp_array performed by
 * orc_interpolate_arrays (src/few/cutils/interpolate.cu:284-325) for fixed
 * knots: given G = dL/d(interp_array) in its (4, ninterps, length) layout,
 * writes dL/dy (ninterps, length).  Steps in reverse order:
 *   final derivatives (:242-280) fold into interval length-2;
 *   fill_coefficients (:28-43): c2 = 3 slope/dx - 2 z_i/dx - z_{i+1}/dx,
 *     c3 = (z_i + z_{i+1} - 2 slope)/dx^2, c1 = z_i, slope = dy/dx;
 *   the not-a-knot solve A z = b (:201-239) -> A^T lambda = dL/dz, solved
 *     with the same restated dgtsv on the transposed tridiagonal;
 *   fill_B (:49-145): b is linear in y with the coefficients of
 *     orc_prep_row, applied transposed.
 * Returns the first non-zero dgtsv INFO (a singular A, e.g. length 3).
 */
int orc_spline_adjoint(const double *t_arr, const double *g_coef, int ninterps, int length,
                       double *g_y)
{
    if (ninterps <= 0 || length < 3) return -1;
    const int L = length;
    const size_t nb = (size_t)ninterps * L;
    double *G = (double *)malloc(sizeof(double) * 4 * (size_t)L);
    double *gz = (double *)malloc(sizeof(double) * (size_t)L);
    double *dl = (double *)malloc(sizeof(double) * (size_t)L);
    double *dd = (double *)malloc(sizeof(double) * (size_t)L);
    double *du = (double *)malloc(sizeof(double) * (size_t)L);
    double *b = (double *)malloc(sizeof(double) * (size_t)L);
    double *ud = (double *)malloc(sizeof(double) * (size_t)L);
    double *diag = (double *)malloc(sizeof(double) * (size_t)L);
    double *ld = (double *)malloc(sizeof(double) * (size_t)L);
    double *zero = (double *)calloc((size_t)L, sizeof(double));
    int first_info = 0;
    if (!G || !gz || !dl || !dd || !du || !b || !ud || !diag || !ld || !zero) {
        first_info = -2;
        goto done;
    }
    for (int s = 0; s < ninterps; s++) {
        const double *x = &t_arr[(size_t)s * L];
        double *gy = &g_y[(size_t)s * L];
        for (int p = 0; p < 4; p++)
            for (int i = 0; i < L; i++) G[p * L + i] = g_coef[p * nb + (size_t)s * L + i];
        /* final derivatives at L-1 */
        {
            const double h = x[L - 1] - x[L - 2];
            const double g1 = G[1 * L + L - 1], g2 = G[2 * L + L - 1], g3 = G[3 * L + L - 1];
            G[1 * L + L - 2] += g1;
            G[2 * L + L - 2] += 2.0 * h * g1 + g2;
            G[3 * L + L - 2] += 3.0 * h * h * g1 + 3.0 * h * g2 + g3;
        }
        for (int i = 0; i < L; i++) {
            gy[i] = G[i];  /* slot 0 is y itself */
            gz[i] = 0.0;
        }
        for (int i = 0; i < L - 1; i++) {
            const double dx = x[i + 1] - x[i];
            const double g1 = G[1 * L + i], g2 = G[2 * L + i], g3 = G[3 * L + i];
            const double gs = 3.0 * g2 / dx - 2.0 * g3 / (dx * dx);
            gz[i] += g1 - 2.0 * g2 / dx + g3 / (dx * dx);
            gz[i + 1] += -g2 / dx + g3 / (dx * dx);
            gy[i] -= gs / dx;
            gy[i + 1] += gs / dx;
        }
        /* the system matrix (orc_prep_row with any y: A depends on x only) */
        for (int i = 0; i < L; i++) orc_prep_row(i, L, b, ud, diag, ld, x, zero);
        /* A^T: sub-diagonal A[i][i+1] = ud[i], super-diagonal A[i+1][i] = ld[i+1] */
        for (int i = 0; i < L - 1; i++) {
            dl[i] = ud[i];
            du[i] = ld[i + 1];
        }
        for (int i = 0; i < L; i++) dd[i] = diag[i];
        int info = orc_dgtsv(L, dl, dd, du, gz);
        if (info != 0 && first_info == 0) first_info = info;
        /* b = B(y): transpose of the rows of orc_prep_row */
        {
            const double dx1 = x[1] - x[0], dx2 = x[2] - x[1], d = x[2] - x[0];
            const double al = (dx1 + 2 * d) * dx2 / (dx1 * d), be = dx1 * dx1 / (dx2 * d);
            gy[0] -= al * gz[0];
            gy[1] += (al - be) * gz[0];
            gy[2] += be * gz[0];
        }
        for (int i = 1; i < L - 1; i++) {
            const double dx1 = x[i] - x[i - 1], dx2 = x[i + 1] - x[i];
            const double al = 3.0 * dx2 / dx1, be = 3.0 * dx1 / dx2;
            gy[i - 1] -= al * gz[i];
            gy[i] += (al - be) * gz[i];
            gy[i + 1] += be * gz[i];
        }
        {
            const double dx1 = x[L - 2] - x[L - 3], dx2 = x[L - 1] - x[L - 2];
            const double d = x[L - 1] - x[L - 3];
            const double al = dx2 * dx2 / (dx1 * d), be = (2 * d + dx2) * dx1 / (dx2 * d);
            gy[L - 3] -= al * gz[L - 1];
            gy[L - 2] += (al - be) * gz[L - 1];
            gy[L - 1] += be * gz[L - 1];
        }
    }
done:
    free(G); free(gz); free(dl); free(dd); free(du); free(b); free(ud); free(diag); free(ld);
    free(zero);
    return first_info;
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

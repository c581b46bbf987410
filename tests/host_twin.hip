// Host twin of the HIP kernels' scalar arithmetic (test infrastructure).
//
// Compiles the product's kernel sources and runs their host-callable
// per-system / per-sample routines on the CPU, so the CPU test suite can
// check the exact code the GPU runs (spline recurrence, dense-output phase,
// the direct mode term, the interval planning rule) against the oracle
// without a GPU.  Only the orchestration (tiling, LDS staging, launch
// geometry) is left to the GPU tests.

#include <stdarg.h>
#include <stdio.h>

#include <vector>

#include "spline.hip"
#include "adjoint.hip"
#include "fd.hip"
#include "amp2d.hip"
#include "aak.hip"
#include "modesum_direct.hpp"
#include "few_math.hpp"

namespace few {
static char g_twin_err[512];
void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_twin_err, sizeof(g_twin_err), fmt, ap);
    va_end(ap);
}
int check_hip(hipError_t e, const char *what)
{
    (void)e;
    set_error("%s", what);
    return FEW_ERR_HIP;
}
// the included launchers reference it; none of them runs on the host
int ensure_dynamic_lds(const void *, int) { return FEW_OK; }
}  // namespace few

using namespace few;

extern "C" {

int twin_spline_build(const double *y, long y_sys, long y_knot, const double *t, long t_sys,
                      long t_knot, double *out, long o_coef, long o_sys, long o_knot, int nsys,
                      int length, int dense, double *scratch, int *info)
{
    SplineGroup g{y, y_sys, y_knot, t, t_sys, t_knot, out, o_coef, o_sys, o_knot, nsys, dense};
    const long f = (long)length * nsys;
    double *sc[4] = {scratch, scratch + f, scratch + 2 * f, scratch + 3 * f};
    SplineArgs a;
    fill_spline_args(&g, 1, length, sc, info, a);
    for (int s = 0; s < nsys; s++) spline_one_system(a, 0, s);
    return 0;
}

int twin_prepare_modes(const double *teuk, const double *phase_knots, const double *t, int length,
                       int num_modes, double *interp_out, double *pcoef_out, double *scratch)
{
    const long K = num_modes, L = length;
    SplineGroup g[3];
    g[0] = SplineGroup{teuk, 2, 2 * K, t, 0, 1, interp_out, L * 2 * K, 1, 2 * K, num_modes, 0};
    g[1] = SplineGroup{teuk + 1, 2, 2 * K, t, 0, 1, interp_out + K, L * 2 * K, 1, 2 * K, num_modes, 0};
    g[2] = SplineGroup{phase_knots, L, 1, t, 0, 1, pcoef_out, 0, 0, 0, 3, 1};
    const long f = L * (2 * K + 3);
    double *sc[4] = {scratch, scratch + f, scratch + 2 * f, scratch + 3 * f};
    SplineArgs a;
    fill_spline_args(g, 3, length, sc, nullptr, a);
    for (int gi = 0; gi < 3; gi++)
        for (int s = 0; s < g[gi].nsys; s++) spline_one_system(a, gi, s);
    return 0;
}

// The shared-knot path of the spline build (spline_factor + spline_system_shared,
// as the batched build runs them per waveform), system by system on the host.
int twin_prepare_modes_shared(const double *teuk, const double *phase_knots, const double *t,
                              int length, int num_modes, double *interp_out, double *pcoef_out)
{
    const long K = num_modes, L = length;
    SplineGroup g[3];
    g[0] = SplineGroup{teuk, 2, 2 * K, t, 0, 1, interp_out, L * 2 * K, 1, 2 * K, num_modes, 0};
    g[1] = SplineGroup{teuk + 1, 2, 2 * K, t, 0, 1, interp_out + K, L * 2 * K, 1, 2 * K, num_modes, 0};
    g[2] = SplineGroup{phase_knots, L, 1, t, 0, 1, pcoef_out, 0, 0, 0, 3, 1};
    std::vector<double> F(5 * (size_t)L), b((size_t)L);
    const int info = spline_factor(t, length, F.data());
    if (info) return info;
    for (int gi = 0; gi < 3; gi++)
        for (int s = 0; s < g[gi].nsys; s++)
            spline_system_shared<8>(g[gi], s, length, t, F.data(), g[gi].y + (long)s * g[gi].y_sys,
                                 g[gi].y_knot, b.data(), 1);
    return 0;
}

// Interval planning exactly as plan_segments_kernel computes it.
int twin_plan(const double *h_t, int L, double dt, int out_len, int *starts, int *ends)
{
    const double T = (out_len - 1) * dt;
    int used = L;
    for (int i = 1; i < L; i++)
        if (!(h_t[i] < T)) {
            used = i + 1;
            break;
        }
    for (int j = 0; j < L - 1; j++) plan_interval(h_t, j, used, L, T, dt, out_len, starts + j, ends + j);
    return used;
}

// Direct FP64 mode sum at chosen samples with the kernel's per-mode code.
int twin_direct_samples(double *out, const long long *idx, long nidx, const double *interp,
                        const double *pt, const double *pcoef, const int *m, const int *k,
                        const int *n, int L, int out_len, int K, const double *ylm, double dt,
                        const double *h_t)
{
    std::vector<int> st(L), en(L);
    twin_plan(h_t, L, dt, out_len, st.data(), en.data());
    const long np = (long)(L - 1) * 3;
    const long nb = (long)L * 2 * K;
    for (long q = 0; q < nidx; q++) {
        const int i = (int)idx[q];
        int j = -1;
        for (int jj = 0; jj < L - 1; jj++)
            if (i >= st[jj] && i < en[jj]) {
                j = jj;
                break;
            }
        double w_re = 0.0, w_im = 0.0;
        if (j >= 0) {
            double rp[8], rt[8], rr[8];
            for (int c = 0; c < 8; c++) {
                rp[c] = pcoef[c * np + (long)j * 3 + 0];
                rt[c] = pcoef[c * np + (long)j * 3 + 1];
                rr[c] = pcoef[c * np + (long)j * 3 + 2];
            }
            SampleState ss;
            sample_state(dt, i, h_t[j], pt[j], pt[j + 1] - pt[j], rp, rt, rr, ss);
            double tr_re = 0.0, tr_im = 0.0;
            for (int mode = 0; mode < K; mode++) {
                if (mode > 0 && mode % 5000 == 0) {
                    w_re = w_re + tr_re;
                    w_im = w_im + tr_im;
                    tr_re = tr_im = 0.0;
                }
                const long br = (long)j * 2 * K + mode, bi = br + K;
                ModeCoef c;
                c.ry = interp[br]; c.rc1 = interp[nb + br]; c.rc2 = interp[2 * nb + br]; c.rc3 = interp[3 * nb + br];
                c.iy = interp[bi]; c.ic1 = interp[nb + bi]; c.ic2 = interp[2 * nb + bi]; c.ic3 = interp[3 * nb + bi];
                c.yp_re = ylm[2 * mode]; c.yp_im = ylm[2 * mode + 1];
                c.ym_re = ylm[2 * (K + mode)]; c.ym_im = ylm[2 * (K + mode) + 1];
                c.m = m[mode]; c.k = k[mode]; c.n = n[mode];
                mode_term_fp64(c, ss.x, ss.x2, ss.x3, ss.Pp, ss.Pt, ss.Pr, tr_re, tr_im);
            }
            w_re = w_re + tr_re;
            w_im = w_im + tr_im;
        }
        out[2 * q] = w_re;
        out[2 * q + 1] = w_im;
    }
    return 0;
}

// The transposed spline build, system by system, on the reference layout
// g_coef (4, nsys, L) -> g_y (nsys, L).  Returns the first non-zero INFO.
int twin_spline_adjoint(const double *t, const double *g_coef, int nsys, int L, double *g_y)
{
    std::vector<double> F(7 * (size_t)L), w((size_t)L);
    const int info = few::tri_factor_transposed(t, L, F.data());
    if (info) return info;
    for (int s = 0; s < nsys; s++)
        few::spline_adjoint_rhs(t, L, F.data(), g_coef + (long)s * L, (long)nsys * L, 1,
                                g_y + (long)s * L, 1, w.data(), 1);
    return 0;
}

void twin_sincos_cw(const double *x, double *s, double *c, int *ok, long n)
{
    for (long i = 0; i < n; i++) ok[i] = sincos_cw(x[i], s[i], c[i]) ? 1 : 0;
}

// The frequency-domain sum with the kernel's per-item and per-bin routines,
// in the CPU branch's loop order (modes, segments, bins, roots) with plain
// sequential adds: bit-comparable with the oracle.
int twin_fd_waveform(double *wave, const double *interp, const double *phase_t,
                     const double *pcoef, const int *m, const int *k, const int *n, int K,
                     const double *t_knots, int L, int Nf, const double *freqs,
                     const int *seg_start, const int *seg_end, int nseg, const double *ylm,
                     int zero_index, int include_minus_m, int separate_modes)
{
    FdArgs a;
    a.wave = reinterpret_cast<double2 *>(wave);
    a.interp = interp;
    a.phase_t = phase_t;
    a.pcoef = pcoef;
    a.m = m;
    a.k = k;
    a.n = n;
    a.t_knots = t_knots;
    a.freqs = freqs;
    a.seg_start = seg_start;
    a.seg_end = seg_end;
    a.ylm = reinterpret_cast<const double2 *>(ylm);
    a.prefix = nullptr;
    a.K = K;
    a.L = L;
    a.nseg = nseg;
    a.data_length = Nf;
    a.zero_index = zero_index;
    a.include_minus_m = include_minus_m && !separate_modes;
    a.separate_modes = separate_modes;
    const fcx xi = fd_xi();
    const FdSpaConst kc = fd_spa_const();
    for (int q = 0; q < K; q++) {
        for (int j = 0; j < nseg; j++) {
            const long item = (long)q * nseg + j;
            int lo, hi;
            fd_item_range(a, item, lo, hi);
            if (hi < lo) continue;
            FdItem it;
            fd_load_item(a, q, j, it);
            for (int i = lo; i <= hi; i++) {
                const int mi = fd_minus_index(zero_index, i);
                const bool conj_term = (it.mm != 0.0) && a.include_minus_m;
                fd_bin(it, freqs[i], xi, kc, [&](fcx plus, fcx minus) {
                    if (separate_modes) {
                        wave[2 * ((long)q * Nf + i)] = plus.re;
                        wave[2 * ((long)q * Nf + i) + 1] = plus.im;
                        return;
                    }
                    wave[2L * i] += plus.re;
                    wave[2L * i + 1] += plus.im;
                    if (conj_term && mi >= 0 && mi < Nf) {
                        wave[2L * mi] += minus.re;
                        wave[2L * mi + 1] += minus.im;
                    }
                });
            }
        }
    }
    return 0;
}

// interp2D with the kernels' per-point basis and per-set sum, in the CPU
// branch's order (sets outer, points inner).
int twin_interp2d(double *z, const double *tx, int nx, const double *ty, int ny, const double *c,
                  int kx, int ky, const double *x, int mx, const double *y, int num_c, int len_c)
{
    std::vector<double> b(14);
    for (int ci = 0; ci < num_c; ci++)
        for (int i = 0; i < mx; i++) {
            amp_point_basis(tx, nx, ty, ny, kx, ky, x[i], y[i], b.data());
            z[(long)ci * mx + i] = amp_sum(c + (long)ci * len_c, ny - ky - 1, kx, ky, b.data());
        }
    return 0;
}

// The AAK sum with the kernel's plan and per-sample routine, segment by
// segment in the CPU branch's order.
int twin_aak_waveform(double *wave, const double *interp, double M, double S, double mu,
                      double qS, double phiS, double qK, double phiK, double dist, int nmodes,
                      int mich, int L, int out_len, double dt, const double *h_t)
{
    AakArgs a;
    a.wave = reinterpret_cast<double2 *>(wave);
    a.interp = interp;
    a.h_t = h_t;
    a.M_phys = M;
    a.S_phys = S;
    a.mu = mu;
    a.qS = qS;
    a.phiS = phiS;
    a.qK = qK;
    a.phiK = phiK;
    a.dist = dist;
    a.dt = dt;
    a.nmodes = nmodes;
    a.mich = mich;
    a.L = L;
    a.out_len = out_len;
    std::vector<int> st((size_t)L + 1);
    a.plan = st.data();
    const int npts = aak_plan(h_t, L, dt, out_len, st.data());
    const long np = (long)(L - 1) * 6;
    for (int j = 0; j < npts - 1; j++) {
        double sc[6][8];
        for (int c = 0; c < 8; c++)
            for (int p = 0; p < 6; p++) sc[p][c] = interp[c * np + (long)j * 6 + p];
        for (long i = st[j]; i < st[j + 1] && i < out_len; i++)
            a.wave[i] = aak_sample(a, sc, h_t[j], h_t[j + 1] - h_t[j], i);
    }
    return 0;
}

// The kernel's Bessel values: nr_jn(order, x) into out1 and the lockstep
// nr_jn5(order, x) (orders order..order+4) into out5, per (order, x) pair.
void twin_bessel(const int *order, const double *x, long n, double *out1, double *out5)
{
    for (long i = 0; i < n; i++) {
        out1[i] = nr_jn(order[i], x[i]);
        double v[5];
        nr_jn5(order[i], x[i], v);
        for (int j = 0; j < 5; j++) out5[5 * i + j] = v[j];
    }
}

void twin_fd_cube_roots(const double *abcd, long n, double *out)
{
    const fcx xi = fd_xi();
    for (long i = 0; i < n; i++)
        fd_cube_roots(abcd[4 * i], abcd[4 * i + 1], abcd[4 * i + 2], abcd[4 * i + 3], xi, out + 3 * i);
}

void twin_fd_spa(const double *x, long n, double *out)
{
    const FdSpaConst kc = fd_spa_const();
    for (long i = 0; i < n; i++) {
        const fcx r = fd_spa(x[i], kc);
        out[2 * i] = r.re;
        out[2 * i + 1] = r.im;
    }
}

// The shared-knot block's pieces against the plain routines, for knots x and
// one array y (n knots): F_plain / F_ahead (5n each: piv | fact | d | du | dl)
// from dgtsv_factor and dgtsv_factor_ahead on the same matrix, their INFOs;
// slopes (n) from dgtsv_sweeps on the right-hand side; c1 (n - 1) from
// spline_system_shared (the slopes it emits).
void twin_spline_block_pieces(const double *x, const double *y, int n, double *F_plain,
                              double *F_ahead, int *info, double *slopes, double *c1)
{
    info[0] = spline_factor(x, n, F_plain);
    for (int r = 0; r < n; r++) {
        const int i0 = (r == 0) ? 0 : (r == n - 1 ? n - 3 : r - 1);
        double b_, u_, g_, l_;
        prep_row(r, n, x[i0], x[i0 + 1], x[i0 + 2], 0.0, 0.0, 0.0, b_, u_, g_, l_);
        F_ahead[2 * n + r] = g_;
        if (r < n - 1) F_ahead[3 * n + r] = u_;
        if (r > 0) F_ahead[4 * n + r - 1] = l_;
    }
    info[1] = dgtsv_factor_ahead<8>(n, F_ahead + 2 * n, F_ahead + 3 * n, F_ahead + 4 * n, F_ahead,
                                    F_ahead + n);
    for (int r = 0; r < n; r++) {  // prep_row's right-hand sides
        const int i0 = (r == 0) ? 0 : (r == n - 1 ? n - 3 : r - 1);
        double b_, u_, g_, l_;
        prep_row(r, n, x[i0], x[i0 + 1], x[i0 + 2], y[i0], y[i0 + 1], y[i0 + 2], b_, u_, g_, l_);
        slopes[r] = b_;
    }
    if (info[1] == 0) dgtsv_sweeps(n, F_ahead, slopes, 1);
    std::vector<double> out(4 * (size_t)n), b(n);
    SplineGroup G{y, 0, 1, x, 0, 1, out.data(), n, 0, 1, 1, 0};
    if (info[0] == 0) spline_system_shared<4>(G, 0, n, x, F_plain, y, 1, b.data(), 1);
    for (int i = 0; i < n - 1; i++) c1[i] = out[n + i];
}

// dgtsv_factor and dgtsv_factor_ahead<8> on copies of one matrix:
// out[k] = piv | fact | d | du | dl (5n) of variant k, info[k].
void twin_dgtsv_pair(int n, const double *d, const double *du, const double *dl, double *out,
                     int *info)
{
    for (int k = 0; k < 2; k++) {
        double *F = out + 5 * (long)n * k;
        for (int i = 0; i < 5 * n; i++) F[i] = 0.0;
        for (int i = 0; i < n; i++) F[2 * n + i] = d[i];
        for (int i = 0; i < n - 1; i++) {
            F[3 * n + i] = du[i];
            F[4 * n + i] = dl[i];
        }
        info[k] = k == 0 ? dgtsv_factor(n, F + 2 * n, F + 3 * n, F + 4 * n, F, F + n)
                         : dgtsv_factor_ahead<8>(n, F + 2 * n, F + 3 * n, F + 4 * n, F, F + n);
    }
}

void twin_sincos_circle(const double *f, double *s, double *c, long n)
{
    for (long i = 0; i < n; i++) sincos_turns_circle(reduce_turns(f[i]), s[i], c[i]);
}

void twin_sincos_turns(const double *f, double *s, double *c, long n)
{
    for (long i = 0; i < n; i++) sincos_turns(reduce_turns(f[i]), s[i], c[i]);
}

}  // extern "C"

// Per-sample / per-mode arithmetic of the direct mode sum, shared by the
// kernels and (host-callable) by the CPU test harness.
#pragma once

#include <math.h>

#include "few_common.hpp"
#include "few_math.hpp"

#ifndef DIRECT_SPT
#define DIRECT_SPT 2  // output samples per lane in the FP64 direct kernel
#endif

namespace few {

struct PlanArgs {
    const double *h_t;   // (L,) amplitude knot times (device)
    const double *pt;    // (L,) phase dense-output segment times (device)
    int L;
    double dt;
    int out_len;
    int tile;            // samples per tile of the consuming kernel
    Segment *segs;       // (L-1,) out
    int *prefix;         // (L,) out: exclusive tile prefix, prefix[L-1] = total
};

struct DirectArgs {
    double2 *wave;            // (out_len,) complex128, accumulated into
    const double *interp;     // (4, L, 2K) knot-major coefficients (reference layout)
    const double *pcoef;      // (8, L-1, 3) dense-output phase coefficients
    const int *m, *k, *n;     // (K,)
    const double2 *ylm;       // (2K,) Y+ then Y- (already times (-1)^l)
    const Segment *segs;      // (L-1,)
    const int *prefix;        // (L,)
    int nseg;                 // L-1
    int L;
    int K;
    double dt;
};

// Output sample index -> start of its interval, find_start_inds rule
// (interpolate.cu:706-717): ceil(t/dt) while the knot is before
// (out_len-1)*dt, out_len afterwards.
__host__ __device__ inline int plan_start(const double *h_t, int j, double T, double dt, int out_len)
{
    if (j == 0) return 0;
    const double t = h_t[j];
    return (t < T) ? (int)ceil(t / dt) : out_len;
}

// Sample range [*st, *en) of interval j, given `used` (the find_start_inds
// stop: the first i >= 1 with h_t[i] >= T gives used = i + 1, else L).  The
// last used interval runs to out_len when the rule stopped early
// (interpolate.cu:711-715).  Intervals at or after used - 1 own no samples
// and are parked at out_len, so the covered samples are [0, en) of the last
// interval in every case.
__host__ __device__ inline void plan_interval(const double *h_t, int j, int used, int L, double T,
                                              double dt, int out_len, int *st, int *en)
{
    int s = out_len, e = out_len;
    if (j < used - 1) {
        s = plan_start(h_t, j, T, dt, out_len);
        e = plan_start(h_t, j + 1, T, dt, out_len);
        if (j + 1 == used - 1 && used < L) e = out_len;
        s = s < 0 ? 0 : (s > out_len ? out_len : s);
        e = e < 0 ? 0 : (e > out_len ? out_len : e);
        if (e < s) e = s;
    }
    *st = s;
    *en = e;
}

struct SampleState {
    double x, x2, x3;     // x = t - h_t[j] and its powers (amplitude cubic)
    double Pp, Pt, Pr;    // Phi_phi, Phi_theta, Phi_r at t
};

// Dense-output phase (interpolate.cu:642-644), contraction off.
__host__ __device__ inline double dense_phase(const double *r, double s, double s1)
{
#pragma clang fp contract(off)
    return r[0] + s * (r[1] + s1 * (r[2] + s * (r[3] + s1 * (r[4] + s * (r[5] + s1 * (r[6] + s * r[7]))))));
}

// The same polynomial with fused steps (seven FMAs instead of seven
// multiplications and seven additions), for the FP16-split paths: its
// rounding differs from dense_phase's by a few ulp of the phase (~1e-10 rad
// at 1e6 rad), far below those paths' ~1e-7 relative error.
__host__ __device__ inline double dense_phase_fused(const double *r, double s, double s1)
{
    double v = fma(s, r[7], r[6]);
    v = fma(s1, v, r[5]);
    v = fma(s, v, r[4]);
    v = fma(s1, v, r[3]);
    v = fma(s, v, r[2]);
    v = fma(s1, v, r[1]);
    return fma(s, v, r[0]);
}

// Sample-level quantities of make_waveform (interpolate.cu:626-644).
__host__ __device__ inline void sample_state(double dt, int i, double h_t, double pt, double pw,
                                             const double *rp, const double *rt, const double *rr,
                                             SampleState &st)
{
#pragma clang fp contract(off)
    const double t = dt * (double)i;
    st.x = t - h_t;
    st.x2 = st.x * st.x;
    st.x3 = st.x * st.x2;
    const double s = (t - pt) / pw;
    const double s1 = 1.0 - s;
    st.Pp = dense_phase(rp, s, s1);
    st.Pt = dense_phase(rt, s, s1);
    st.Pr = dense_phase(rr, s, s1);
}

struct ModeCoef {
    double ry, rc1, rc2, rc3;     // real-part cubic
    double iy, ic1, ic2, ic3;     // imaginary-part cubic
    double yp_re, yp_im, ym_re, ym_im;
    double m, k, n;
};

// m*Phi_phi + k*Phi_theta + n*Phi_r rounded exactly as the oracle
// (interpolate.cu:662); contraction off.
__host__ __device__ inline double mode_phase(double m, double Pp, double k, double Pt, double n,
                                             double Pr)
{
#pragma clang fp contract(off)
    return m * Pp + k * Pt + n * Pr;
}

// One mode's contribution at one sample (interpolate.cu:647-684):
//   P = exp(-i*phase) * A(x);  tr += [m != 0] conj(P) Y- ;  tr += P Y+
__host__ __device__ inline void mode_term_fp64(const ModeCoef &c, double x, double x2, double x3,
                                               double Pp, double Pt, double Pr, double &tr_re,
                                               double &tr_im)
{
    const double phase = mode_phase(c.m, Pp, c.k, Pt, c.n, Pr);
    double e_im, e_re;
    if (!sincos_cw(-phase, e_im, e_re)) sincos(-phase, &e_im, &e_re);
    const double a_re = c.ry + c.rc1 * x + c.rc2 * x2 + c.rc3 * x3;
    const double a_im = c.iy + c.ic1 * x + c.ic2 * x2 + c.ic3 * x3;
    const double p_re = e_re * a_re - e_im * a_im;
    const double p_im = e_re * a_im + e_im * a_re;
    const double tp_re = p_re * c.yp_re - p_im * c.yp_im;
    const double tp_im = p_re * c.yp_im + p_im * c.yp_re;
    double tm_re = 0.0, tm_im = 0.0;
    if (c.m != 0.0) {
        tm_re = p_re * c.ym_re + p_im * c.ym_im;
        tm_im = p_re * c.ym_im - p_im * c.ym_re;
    }
    tr_re = tr_re + tm_re + tp_re;
    tr_im = tr_im + tm_im + tp_im;
}

int launch_plan(const PlanArgs &p, hipStream_t stream);
int direct_tile(int precision);
int launch_direct_fp64(const DirectArgs &a, int max_blocks, hipStream_t stream);

}  // namespace few

// Spline evaluation and the dense direct mode sum (eval.hip).
#pragma once

#include "few_common.hpp"

namespace few {

struct SplineEvalArgs {
    const double *t;        // knots: t[s*t_sys + j*t_knot]
    long t_sys, t_knot;
    int length, ninterps;
    const double *coef;     // coefficient p of array s at knot j: coef[p*c_coef + s*c_sys + j*c_knot]
    long c_coef, c_sys, c_knot;
    const double *tnew;     // tnew[s*tn_sys + q]
    long tn_sys;
    int npts;
    int deriv;
    double *out;            // out[s*npts + q]
    int *flags;             // [0] any point left of the range, [1] any point right of it
};

struct DenseSumArgs {
    const double2 *teuk;       // (N, K) complex amplitudes at the output samples
    const double2 *ylm;        // (2K,) Y+ then Y- (as passed to DirectModeSum.sum)
    const double *phi_phi, *phi_theta, *phi_r;   // (N,)
    const int *l, *m, *k, *n;  // (K,)
    int N, K;
    double2 *out;              // (N,) written (not accumulated)
};

int launch_spline_eval(const SplineEvalArgs &a, hipStream_t stream);
int launch_dense_sum(const DenseSumArgs &a, hipStream_t stream);

}  // namespace few

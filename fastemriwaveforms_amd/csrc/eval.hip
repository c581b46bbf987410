// Spline evaluation (CubicSplineInterpolant.__call__) and the direct mode sum
// over dense per-sample amplitudes and phases (DirectModeSum.sum).
//
// Spline evaluation (summation/interpolatedmodesum.py:139-257): one lane per
// (array, point).  The interval is found by binary search with the
// reference's rule: searchsorted(t, x, side="right") - 1, and x == t[-1]
// belongs to the last interval.  The value or derivative is evaluated with
// the reference's expressions and operation order (contraction off), so the
// result equals numpy's bit for bit.  Order 0 fills points outside the knot
// range with the edge values; the flags report points outside the range
// (flags[0] left, flags[1] right) so the caller can warn, or raise for
// derivatives, as the reference does.
//
// Direct mode sum (summation/directmodesum.py:70-126): one lane per output
// sample; mode data staged through LDS in chunks; per mode
//   h += Y+ A e^{-i phi} + [m > 0] (-1)^l Y- conj(A) e^{+i phi},
// phi = m Phi_phi + k Phi_theta + n Phi_r, FP64 throughout.  The amplitude
// tile of a workgroup (its samples x a chunk of modes) is read with the
// modes contiguous, so the (N, K) amplitude array is read coalesced.
#include <math.h>

#include "eval.hpp"
#include "few_common.hpp"
#include "few_math.hpp"
#include "modesum_direct.hpp"

namespace few {

__global__ __launch_bounds__(256) void spline_eval_kernel(SplineEvalArgs a)
{
#pragma clang fp contract(off)
    const long gid = (long)blockIdx.x * 256 + threadIdx.x;
    if (gid >= (long)a.ninterps * a.npts) return;
    const int s = (int)(gid / a.npts);
    const int q = (int)(gid - (long)s * a.npts);
    const double *t = a.t + (long)s * a.t_sys;
    const long tk = a.t_knot;
    const int L = a.length;
    const double x0 = a.tnew[(long)s * a.tn_sys + q];
    const double tfirst = t[0], tlast = t[(long)(L - 1) * tk];
    // largest j with t[j] <= x0 (j = -1 when x0 < t[0]); x0 == t[L-1] -> L-2
    int lo = -1, hi = L;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (t[(long)mid * tk] <= x0) lo = mid; else hi = mid;
    }
    int j = lo;
    if (x0 == tlast) j = L - 2;
    const bool bad_left = x0 < tfirst, bad_right = x0 > tlast;
    if (bad_left) a.flags[0] = 1;    // benign race: every writer stores 1
    if (bad_right) a.flags[1] = 1;
    j = j < 0 ? 0 : (j > L - 1 ? L - 1 : j);
    const double *c = a.coef + (long)s * a.c_sys + (long)j * a.c_knot;
    const double y = c[0], c1 = c[a.c_coef], c2 = c[2 * a.c_coef], c3 = c[3 * a.c_coef];
    const double x = x0 - t[(long)j * tk];
    const double x2 = x * x;
    const double x3 = x2 * x;
    double v;
    if (a.deriv == 0) {
        v = y + c1 * x + c2 * x2 + c3 * x3;
        if (bad_left) v = a.coef[(long)s * a.c_sys];
        if (bad_right) v = a.coef[(long)s * a.c_sys + (long)(L - 1) * a.c_knot];
    } else if (a.deriv == 1) {
        v = c1 + 2 * c2 * x + 3 * c3 * x2;
    } else if (a.deriv == 2) {
        v = 2 * c2 + 6 * c3 * x;
    } else {
        v = 6 * c3;
    }
    a.out[(long)s * a.npts + q] = v;
}

int launch_spline_eval(const SplineEvalArgs &a, hipStream_t stream)
{
    const long n = (long)a.ninterps * a.npts;
    if (n <= 0) return FEW_OK;
    if ((n + 255) / 256 > 0x7fffffffL) {
        set_error("spline_eval: too many points (%ld)", n);
        return FEW_ERR_LIMIT;
    }
    hipLaunchKernelGGL(spline_eval_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                       a);
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

// ------------------------------------------------------------ direct mode sum

#define DS_SAMPLES 64          // samples (lanes) per workgroup
#define DS_MODES 32            // modes staged per step (LDS tile 64 x 33 complex = 33.8 KB)

__global__ __launch_bounds__(DS_SAMPLES) void dense_sum_kernel(DenseSumArgs a)
{
    __shared__ double2 sA[DS_SAMPLES][DS_MODES + 1];
    __shared__ double sY[4][DS_MODES];
    __shared__ double sQ[5][DS_MODES];   // m, k, n, (-1)^l, [m > 0]
    const int tid = threadIdx.x;
    const long i0 = (long)blockIdx.x * DS_SAMPLES;
    const long i = i0 + tid;
    const bool ok = i < a.N;
    double pp = 0.0, pt = 0.0, pr = 0.0;
    if (ok) {
        pp = a.phi_phi[i];
        pt = a.phi_theta[i];
        pr = a.phi_r[i];
    }
    double hr = 0.0, hi = 0.0;
    for (int q0 = 0; q0 < a.K; q0 += DS_MODES) {
        const int cnt = min(DS_MODES, a.K - q0);
        __syncthreads();
        // amplitude tile: row r (sample i0 + r), modes q0 .. q0+cnt-1, read along the modes
        for (int r = 0; r < DS_SAMPLES; r++) {
            const long ii = i0 + r;
            if (ii < a.N && tid < cnt) sA[r][tid] = a.teuk[ii * a.K + q0 + tid];
        }
        if (tid < cnt) {
            const int q = q0 + tid;
            const double2 yp = a.ylm[q], ym = a.ylm[a.K + q];
            sY[0][tid] = yp.x; sY[1][tid] = yp.y; sY[2][tid] = ym.x; sY[3][tid] = ym.y;
            sQ[0][tid] = (double)a.m[q];
            sQ[1][tid] = (double)a.k[q];
            sQ[2][tid] = (double)a.n[q];
            sQ[3][tid] = (a.l[q] & 1) ? -1.0 : 1.0;
            sQ[4][tid] = a.m[q] > 0 ? 1.0 : 0.0;
        }
        __syncthreads();
        if (!ok) continue;
        for (int u = 0; u < cnt; u++) {
            const double phase = mode_phase(sQ[0][u], pp, sQ[1][u], pt, sQ[2][u], pr);
            double sn, cs;
            if (!sincos_cw(phase, sn, cs)) sincos(phase, &sn, &cs);
            const double2 A = sA[tid][u];
            // Y+ A e^{-i phi}
            const double ya_r = sY[0][u] * A.x - sY[1][u] * A.y;
            const double ya_i = sY[0][u] * A.y + sY[1][u] * A.x;
            hr += ya_r * cs + ya_i * sn;
            hi += ya_i * cs - ya_r * sn;
            // [m > 0] (-1)^l Y- conj(A) e^{+i phi}
            const double w = sQ[3][u] * sQ[4][u];
            const double yb_r = w * (sY[2][u] * A.x + sY[3][u] * A.y);
            const double yb_i = w * (sY[3][u] * A.x - sY[2][u] * A.y);
            hr += yb_r * cs - yb_i * sn;
            hi += yb_i * cs + yb_r * sn;
        }
    }
    if (ok) a.out[i] = make_double2(hr, hi);
}

int launch_dense_sum(const DenseSumArgs &a, hipStream_t stream)
{
    if (a.N <= 0) return FEW_OK;
    const long blocks = ((long)a.N + DS_SAMPLES - 1) / DS_SAMPLES;
    if (blocks > 0x7fffffffL) {
        set_error("direct_mode_sum: too many samples (%d)", a.N);
        return FEW_ERR_LIMIT;
    }
    hipLaunchKernelGGL(dense_sum_kernel, dim3((unsigned)blocks), dim3(DS_SAMPLES), 0, stream, a);
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

}  // namespace few

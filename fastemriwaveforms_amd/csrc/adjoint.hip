// Reverse-mode kernels of the interpolated mode sum (FP64) and of the
// not-a-knot spline build.  The formulas are stated in adjoint.hpp.
//
// Mode-sum adjoint: one workgroup owns one chunk of one interval's samples
// (at most `chunk` samples, never crossing an interval boundary) and one
// group of ADJ_MODES modes, one mode per lane.  The chunk's sample state
// (x, the three phases, g) is computed once per sample into LDS, then every
// lane walks the staged samples and keeps its mode's twelve moments in
// registers: sum x^p dL/dRe A, sum x^p dL/dIm A (p = 0..3) and
// sum xp^p dL/dphi.  Nothing is reduced across lanes per sample; per chunk
// the amplitude moments go to a partial buffer (chunk, 4, 2K) and the phase
// moments (weighted by m, k, n) are reduced over the workgroup.  A second
// kernel adds the chunks of each interval in a fixed order (deterministic).
#include <math.h>

#include "adjoint.hpp"
#include "few_math.hpp"
#include "modesum_direct.hpp"
#include "tridiag.hpp"

namespace few {

#define ADJ_MODES 128   // lanes (modes) per workgroup
#define ADJ_SUB 128     // samples staged in LDS per step

// Transposed spline build, in two parts.  Every system of one call shares
// the knots, hence the matrix A^T of the not-a-knot system; LAPACK dgtsv's
// pivot choices and multipliers depend on the matrix only, so the
// elimination is done once (tri_factor_transposed) and each system then
// sweeps only its right-hand side with the recorded multipliers
// (spline_adjoint_rhs).  Every operation is the one dgtsv / the transposed
// build performs, in the same order, with contraction off: the result equals
// the one-system-at-a-time transpose bit for bit.
//
// F (7 x L doubles): [0] pivot flag of elimination step i, [1] its
// multiplier, [2] d, [3] du, [4] dl (second super-diagonal fill-in) after
// elimination, [5] al and [6] be, the weights of row i of fill_B transposed.
__host__ __device__ int tri_factor_transposed(const double *x, int L, double *F)
{
#pragma clang fp contract(off)
    double *piv = F, *fact = F + L, *d = F + 2L * L, *du = F + 3L * L, *dl = F + 4L * L;
    double *al = F + 5L * L, *be = F + 6L * L;
    // A^T of prep_splines' matrix (interpolate.cu:49-102): sub-diagonal
    // A[i][i+1], super-diagonal A[i+1][i]
    for (int i = 0; i < L; i++) {
        double sub = 0.0, sup = 0.0, dg;
        if (i == L - 1) {
            dg = x[L - 2] - x[L - 3];
        } else if (i == 0) {
            dg = x[2] - x[1];
            sub = x[2] - x[0];
        } else {
            dg = 2 * ((x[i] - x[i - 1]) + (x[i + 1] - x[i]));
            sub = x[i] - x[i - 1];
        }
        if (i < L - 1) sup = (i + 1 == L - 1) ? x[L - 1] - x[L - 3] : x[i + 2] - x[i + 1];
        d[i] = dg;
        dl[i] = sub;
        du[i] = sup;
    }
    const int info = dgtsv_factor(L, d, du, dl, piv, fact);
    if (info) return info;
    // fill_B (interpolate.cu:107-145): weights of row i on y[i-1] (-al), y[i] (al - be), y[i+1] (be)
    {
        const double dx1 = x[1] - x[0], dx2 = x[2] - x[1], dd = x[2] - x[0];
        al[0] = (dx1 + 2 * dd) * dx2 / (dx1 * dd);
        be[0] = dx1 * dx1 / (dx2 * dd);
    }
    for (int r = 1; r < L - 1; r++) {
        const double dx1 = x[r] - x[r - 1], dx2 = x[r + 1] - x[r];
        al[r] = 3.0 * dx2 / dx1;
        be[r] = 3.0 * dx1 / dx2;
    }
    {
        const double dx1 = x[L - 2] - x[L - 3], dx2 = x[L - 1] - x[L - 2];
        const double dd = x[L - 1] - x[L - 3];
        al[L - 1] = dx2 * dx2 / (dx1 * dd);
        be[L - 1] = (2 * dd + dx2) * dx1 / (dx2 * dd);
    }
    return 0;
}

// One system: gradient G(p, i) = g[p*gp + i*gk] of its (4, L) coefficients ->
// dL/dy(i) = gy[i*gyk].  s: L scratch doubles with stride ws.  Two passes:
// spline_adjoint_pass1 needs only the knots (fill_coefficients and the final
// derivatives transposed); spline_adjoint_solve needs the elimination F.
__host__ __device__ __attribute__((always_inline)) void spline_adjoint_pass1(
    const double *x, int L, const double *g, long gp, long gk, double *gy, long gyk, double *s,
    long ws)
{
#pragma clang fp contract(off)
#define G(p, i) g[(long)(p) * gp + (long)(i) * gk]
#define GY(i) gy[(long)(i) * gyk]
#define S(i) s[(long)(i) * ws]
    // final derivatives (interpolate.cu:242-280) fold into interval L-2;
    // fill_coefficients (interpolate.cu:28-43) transposed, interval by
    // interval: z = dL/d(slopes) into s, the direct y terms into gy
    const double h = x[L - 1] - x[L - 2];
    const double f1 = G(1, L - 1), f2 = G(2, L - 1), f3 = G(3, L - 1);
    double pend_z = 0.0, q_prev = 0.0;
    auto interval = [&](int i, double g0, double g1, double g2, double g3) {
        const double dx = x[i + 1] - x[i];
        if (i == L - 2) {
            g1 += f1;
            g2 += 2.0 * h * f1 + f2;
            g3 += 3.0 * h * h * f1 + 3.0 * h * f2 + f3;
        }
        const double gs = 3.0 * g2 / dx - 2.0 * g3 / (dx * dx);
        S(i) = pend_z + (g1 - 2.0 * g2 / dx + g3 / (dx * dx));
        pend_z = 0.0 + (-g2 / dx + g3 / (dx * dx));
        const double q = gs / dx;
        GY(i) = (i > 0 ? g0 + q_prev : g0) - q;
        q_prev = q;
    };
    // The gradient loads are issued ADJ_PF intervals ahead through a register
    // ring so the sequential pass does not wait on memory each interval.
    int i = 0;
    if (L - 1 >= ADJ_PF) {
        double r0[ADJ_PF], r1[ADJ_PF], r2[ADJ_PF], r3[ADJ_PF];
#pragma unroll
        for (int u = 0; u < ADJ_PF; u++) {
            r0[u] = G(0, u);
            r1[u] = G(1, u);
            r2[u] = G(2, u);
            r3[u] = G(3, u);
        }
        for (; i + ADJ_PF <= L - 1; i += ADJ_PF) {
#pragma unroll
            for (int u = 0; u < ADJ_PF; u++) {
                const double g0 = r0[u], g1 = r1[u], g2 = r2[u], g3 = r3[u];
                // unconditional (clamped) loads keep the ring free of branches,
                // which would make the compiler wait for them at once
                const int nx = min(i + u + ADJ_PF, L - 1);
                r0[u] = G(0, nx);
                r1[u] = G(1, nx);
                r2[u] = G(2, nx);
                r3[u] = G(3, nx);
                interval(i + u, g0, g1, g2, g3);
            }
        }
    }
    for (; i < L - 1; i++) interval(i, G(0, i), G(1, i), G(2, i), G(3, i));
    S(L - 1) = pend_z;
    GY(L - 1) = G(0, L - 1) + q_prev;
#undef G
}

__host__ __device__ __attribute__((always_inline)) void spline_adjoint_solve(
    const double *x, int L, const double *F, double *gy, long gyk, double *s, long ws)
{
#pragma clang fp contract(off)
    const double *piv = F, *fact = F + L, *d = F + 2L * L, *du = F + 3L * L, *dl = F + 4L * L;
    const double *al = F + 5L * L, *be = F + 6L * L;
    // the solve A^T lambda = z: forward sweep with the recorded multipliers
    double cur = S(0);
    for (int i = 0; i < L - 1; i++) {
        const double nxt = S(i + 1);
        if (piv[i] == 0.0) {
            S(i) = cur;
            cur = nxt - fact[i] * cur;
        } else {
            S(i) = nxt;
            cur = cur - fact[i] * nxt;
        }
    }
    // back substitution
    double b1 = cur / d[L - 1];
    S(L - 1) = b1;
    double b0 = (S(L - 2) - du[L - 2] * b1) / d[L - 2];
    S(L - 2) = b0;
    for (int i = L - 3; i >= 0; i--) {
        const double bi = (S(i) - du[i] * b0 - dl[i] * b1) / d[i];
        S(i) = bi;
        b1 = b0;
        b0 = bi;
    }
    // fill_B (interpolate.cu:107-145) transposed, rows in ascending order:
    // row r adds to gy[r-1], gy[r], gy[r+1]; a three-value window
    double y0 = GY(0), y1 = GY(1), y2 = GY(2);
    double lam = S(0);
    y0 -= al[0] * lam;
    y1 += (al[0] - be[0]) * lam;
    y2 += be[0] * lam;
    for (int r = 1; r <= L - 2; r++) {
        lam = S(r);
        y0 -= al[r] * lam;
        y1 += (al[r] - be[r]) * lam;
        y2 += be[r] * lam;
        if (r < L - 2) {
            GY(r - 1) = y0;
            y0 = y1;
            y1 = y2;
            y2 = GY(r + 2);
        }
    }
    lam = S(L - 1);
    y0 -= al[L - 1] * lam;
    y1 += (al[L - 1] - be[L - 1]) * lam;
    y2 += be[L - 1] * lam;
    GY(L - 3) = y0;
    GY(L - 2) = y1;
    GY(L - 1) = y2;
#undef GY
#undef S
}

__host__ __device__ __attribute__((always_inline)) void spline_adjoint_rhs(
    const double *x, int L, const double *F, const double *g, long gp, long gk, double *gy,
    long gyk, double *s, long ws)
{
    spline_adjoint_pass1(x, L, g, gp, gk, gy, gyk, s, ws);
    spline_adjoint_solve(x, L, F, gy, gyk, s, ws);
}

__global__ __launch_bounds__(64) void tri_factor_kernel(const double *t, int L, double *F,
                                                        int *info)
{
    if (threadIdx.x == 0) {
        const int r = tri_factor_transposed(t, L, F);
        if (info) *info = r;
    }
}

// Row r of tri_factor_transposed's matrix (d, dl, du) and of its fill_B
// weights (al, be): the same expressions, one row.
__device__ inline void tri_row_transposed(const double *x, int L, int r, double *F)
{
#pragma clang fp contract(off)
    double *d = F + 2L * L, *du = F + 3L * L, *dl = F + 4L * L;
    double *al = F + 5L * L, *be = F + 6L * L;
    double sub = 0.0, sup = 0.0, dg;
    if (r == L - 1) {
        dg = x[L - 2] - x[L - 3];
    } else if (r == 0) {
        dg = x[2] - x[1];
        sub = x[2] - x[0];
    } else {
        dg = 2 * ((x[r] - x[r - 1]) + (x[r + 1] - x[r]));
        sub = x[r] - x[r - 1];
    }
    if (r < L - 1) sup = (r + 1 == L - 1) ? x[L - 1] - x[L - 3] : x[r + 2] - x[r + 1];
    d[r] = dg;
    dl[r] = sub;
    du[r] = sup;
    if (r == 0) {
        const double dx1 = x[1] - x[0], dx2 = x[2] - x[1], dd = x[2] - x[0];
        al[0] = (dx1 + 2 * dd) * dx2 / (dx1 * dd);
        be[0] = dx1 * dx1 / (dx2 * dd);
    } else if (r == L - 1) {
        const double dx1 = x[L - 2] - x[L - 3], dx2 = x[L - 1] - x[L - 2];
        const double dd = x[L - 1] - x[L - 3];
        al[L - 1] = dx2 * dx2 / (dx1 * dd);
        be[L - 1] = (2 * dd + dx2) * dx1 / (dx2 * dd);
    } else {
        const double dx1 = x[r] - x[r - 1], dx2 = x[r + 1] - x[r];
        al[r] = 3.0 * dx2 / dx1;
        be[r] = 3.0 * dx1 / dx2;
    }
}

// spline_adjoint_pass1's work for interval i of one system (the same
// expressions): z part of S(i), the pending z it hands to S(i+1), q.
struct AdjIv {
    double zpart, pend, q, g0;
};
__device__ inline AdjIv adj_interval(const double *x, int L, const double *g, long gp, long gk,
                                     int i)
{
#pragma clang fp contract(off)
#define G(p, j) g[(long)(p) * gp + (long)(j) * gk]
    double g0 = G(0, i), g1 = G(1, i), g2 = G(2, i), g3 = G(3, i);
    const double dx = x[i + 1] - x[i];
    if (i == L - 2) {
        const double h = x[L - 1] - x[L - 2];
        const double f1 = G(1, L - 1), f2 = G(2, L - 1), f3 = G(3, L - 1);
        g1 += f1;
        g2 += 2.0 * h * f1 + f2;
        g3 += 3.0 * h * h * f1 + 3.0 * h * f2 + f3;
    }
#undef G
    const double gs = 3.0 * g2 / dx - 2.0 * g3 / (dx * dx);
    AdjIv v;
    v.zpart = g1 - 2.0 * g2 / dx + g3 / (dx * dx);
    v.pend = 0.0 + (-g2 / dx + g3 / (dx * dx));
    v.q = gs / dx;
    v.g0 = g0;
    return v;
}

#define ADJ_SPB 32  // systems per shared-knot block (half a wave)

// One workgroup: ADJ_SPB systems of group gi on the shared knots, four
// waves, the knots, the elimination and both right-hand-side tiles
// (scratch z / lambda and the knot-value gradients) in LDS; the arithmetic is
// spline_adjoint_rhs's, bit for bit.  What is parallel over (system, knot)
// is spread over the waves; only the sweeps are sequential along the knots:
//   wave 3:    A^T's rows and fill_B weights (one row per lane), then dgtsv's
//              elimination chain (every lane the same chain);
//   waves 0-2: meanwhile the first pass, one (system, interval) per lane
//              (each lane recomputes the previous interval's carries);
//   wave 0:    the sweeps of A^T lambda = z, one system per lane;
//   waves 0-3: fill_B transposed, one (system, knot) per lane: the rows
//              touching the knot in ascending order, as the sequential pass
//              adds them.
__global__ __launch_bounds__(256) void spline_adjoint_lds_kernel(SplineAdjArgs a)
{
    constexpr int S = ADJ_SPB;
    extern __shared__ __attribute__((aligned(16))) double adj_lds[];
    int gi = 0;
    while (gi + 1 < a.ngroups && (int)blockIdx.x >= a.blk_prefix[gi + 1]) gi++;
    const SplineAdjGroup &G = a.grp[gi];
    const int L = a.L, tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6, half = lane / S, ls = lane % S;
    const long s0 = (long)((int)blockIdx.x - a.blk_prefix[gi]) * S;
    const int nloc = (int)min((long)S, G.nsys - s0);
    double *sx = adj_lds, *F = sx + L, *sS = F + 7L * L, *sG = sS + (long)S * L;
    for (int i = tid; i < L; i += 256) sx[i] = a.t[i];
    __syncthreads();
    if (wv == 3) {
        for (int r = lane; r < L; r += 64) tri_row_transposed(sx, L, r, F);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        (void)dgtsv_factor_ahead<8>(L, F + 2L * L, F + 3L * L, F + 4L * L, F, F + L);
    } else if (ls < nloc) {
        const double *g = G.g + (s0 + ls) * G.g_sys;
        // interval i on lane group 2 wv + half of waves 0-2 (6 groups)
        for (int i = 2 * wv + half; i < L - 1; i += 6) {
            const AdjIv v = adj_interval(sx, L, g, G.g_p, G.g_knot, i);
            double pend_prev = 0.0, q_prev = 0.0;
            if (i > 0) {
                const AdjIv w = adj_interval(sx, L, g, G.g_p, G.g_knot, i - 1);
                pend_prev = w.pend;
                q_prev = w.q;
            }
            sS[(long)i * S + ls] = pend_prev + v.zpart;
            sG[(long)i * S + ls] = (i > 0 ? v.g0 + q_prev : v.g0) - v.q;
            if (i == L - 2) {
                sS[(long)(L - 1) * S + ls] = v.pend;
                sG[(long)(L - 1) * S + ls] = g[(long)(L - 1) * G.g_knot] + v.q;
            }
        }
    }
    __syncthreads();
    if (wv == 0 && half == 0) dgtsv_sweeps(L, F, sS + ls, S);
    __syncthreads();
    if (ls >= nloc) return;
    {
#pragma clang fp contract(off)
        const double *al = F + 5L * L, *be = F + 6L * L;
        double *gy = G.gy + (s0 + ls) * G.gy_sys;
        for (int k = 2 * wv + half; k < L; k += 8) {
            double y = sG[(long)k * S + ls];
            const int r0 = k - 2 > 0 ? k - 2 : 0, r1 = k + 2 < L - 1 ? k + 2 : L - 1;
            for (int r = r0; r <= r1; r++) {
                const int st = (r == 0) ? 0 : (r == L - 1 ? L - 3 : r - 1);
                const int pos = k - st;
                const double lam = sS[(long)r * S + ls];
                if (pos == 0)
                    y -= al[r] * lam;
                else if (pos == 1)
                    y += (al[r] - be[r]) * lam;
                else if (pos == 2)
                    y += be[r] * lam;
            }
            gy[(long)k * G.gy_knot] = y;
        }
    }
}

__global__ __launch_bounds__(64) void spline_adjoint_kernel(SplineAdjArgs a)
{
    long s = (long)blockIdx.x * 64 + threadIdx.x;
    const long col = s;
    for (int gi = 0; gi < a.ngroups; gi++) {
        const SplineAdjGroup &G = a.grp[gi];
        if (s < G.nsys) {
            spline_adjoint_rhs(a.t, a.L, a.factor, G.g + s * G.g_sys, G.g_p, G.g_knot,
                               G.gy + s * G.gy_sys, G.gy_knot, a.scratch + col, a.ns);
            return;
        }
        s -= G.nsys;
    }
}

// ------------------------------------------------------------ mode-sum adjoint

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ __launch_bounds__(ADJ_MODES) void adjoint_fp64_kernel(AdjArgs a)
{
    __shared__ double s_x[ADJ_SUB], s_pp[ADJ_SUB], s_pt[ADJ_SUB], s_pr[ADJ_SUB];
    __shared__ double s_gr[ADJ_SUB], s_gi[ADJ_SUB], s_xp[ADJ_SUB];
    __shared__ double s_red[ADJ_MODES / 64][12];

    const int c = blockIdx.x;
    const int total = a.prefix[a.nseg];
    if (c >= total) return;  // uniform over the workgroup
    int lo_s = 0, hi_s = a.nseg;
    while (hi_s - lo_s > 1) {
        const int mid = (lo_s + hi_s) >> 1;
        if (a.prefix[mid] <= c) lo_s = mid; else hi_s = mid;
    }
    const int j = lo_s;
    const Segment sg = a.segs[j];
    const int lo = sg.start + (c - a.prefix[j]) * a.chunk;
    const int hi = min(lo + a.chunk, sg.end);
    const int tid = threadIdx.x;
    const int K = a.K;
    const int mode = blockIdx.y * ADJ_MODES + tid;
    const bool act = mode < K;

    double cr0 = 0, cr1 = 0, cr2 = 0, cr3 = 0, ci0 = 0, ci1 = 0, ci2 = 0, ci3 = 0;
    double ypr = 0, ypi = 0, ymr = 0, ymi = 0, mm = 0, kk = 0, nn = 0;
    if (act) {
        const long nb = (long)a.L * 2 * K;
        const long br = (long)j * 2 * K + mode, bi = br + K;
        cr0 = a.interp[br]; cr1 = a.interp[nb + br]; cr2 = a.interp[2 * nb + br]; cr3 = a.interp[3 * nb + br];
        ci0 = a.interp[bi]; ci1 = a.interp[nb + bi]; ci2 = a.interp[2 * nb + bi]; ci3 = a.interp[3 * nb + bi];
        const double2 yp = a.ylm[mode];
        ypr = yp.x;
        ypi = yp.y;
        const int m = a.m[mode];
        if (m != 0) {
            const double2 ym = a.ylm[K + mode];
            ymr = ym.x;
            ymi = ym.y;
        }
        mm = (double)m;
        kk = (double)a.k[mode];
        nn = (double)a.n[mode];
    }
    const long np = (long)a.nseg * 3;
    double rp[8], rt[8], rr[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        rp[q] = a.pcoef[q * np + (long)j * 3 + 0];
        rt[q] = a.pcoef[q * np + (long)j * 3 + 1];
        rr[q] = a.pcoef[q * np + (long)j * 3 + 2];
    }

    double dr0 = 0, dr1 = 0, dr2 = 0, dr3 = 0, di0 = 0, di1 = 0, di2 = 0, di3 = 0;
    double R0 = 0, R1 = 0, R2 = 0, R3 = 0;
    for (int base = lo; base < hi; base += ADJ_SUB) {
        __syncthreads();
        {
            const int i = base + tid;
            SampleState st;
            sample_state(a.dt, i, sg.h_t, sg.pt, sg.pw, rp, rt, rr, st);
            double gr = 0.0, gi = 0.0;
            if (i < hi) {
                const double2 g = a.grad[i];
                gr = g.x;
                gi = g.y;
            }
            s_x[tid] = st.x;
            s_pp[tid] = st.Pp;
            s_pt[tid] = st.Pt;
            s_pr[tid] = st.Pr;
            s_gr[tid] = gr;
            s_gi[tid] = gi;
            s_xp[tid] = a.dt * (double)i - sg.pt;
        }
        __syncthreads();
        const int cnt = min(ADJ_SUB, hi - base);
        for (int q = 0; q < cnt; q++) {
            const double x = s_x[q];
            const double phase = mode_phase(mm, s_pp[q], kk, s_pt[q], nn, s_pr[q]);
            double e_im, e_re;
            if (!sincos_cw(-phase, e_im, e_re)) sincos(-phase, &e_im, &e_re);
            const double gr = s_gr[q], gi = s_gi[q];
            // q+ = conj(g) e, q- = conj(g) conj(e)
            const double A1 = gr * e_re, A2 = gi * e_im, A3 = gr * e_im, A4 = gi * e_re;
            const double qpr = A1 + A2, qpi = A3 - A4, qmr = A1 - A2, qmi = -(A3 + A4);
            const double Wr = qpr * ypr - qpi * ypi, Wi = qpr * ypi + qpi * ypr;
            const double Vr = qmr * ymr - qmi * ymi, Vi = qmr * ymi + qmi * ymr;
            const double d_re = Wr + Vr, d_im = Vi - Wi;
            const double a_re = cr0 + x * (cr1 + x * (cr2 + x * cr3));
            const double a_im = ci0 + x * (ci1 + x * (ci2 + x * ci3));
            const double rho = a_im * d_re - a_re * d_im;
            const double x2 = x * x, x3 = x2 * x;
            dr0 += d_re; dr1 += x * d_re; dr2 += x2 * d_re; dr3 += x3 * d_re;
            di0 += d_im; di1 += x * d_im; di2 += x2 * d_im; di3 += x3 * d_im;
            const double xp = s_xp[q], xp2 = xp * xp, xp3 = xp2 * xp;
            R0 += rho; R1 += xp * rho; R2 += xp2 * rho; R3 += xp3 * rho;
        }
    }
    if (act) {
        const long K2 = 2L * K;
        double *pa = a.part_amp + (long)c * 4 * K2;
        pa[mode] = dr0;          pa[K + mode] = di0;
        pa[K2 + mode] = dr1;     pa[K2 + K + mode] = di1;
        pa[2 * K2 + mode] = dr2; pa[2 * K2 + K + mode] = di2;
        pa[3 * K2 + mode] = dr3; pa[3 * K2 + K + mode] = di3;
    }
    // phase moments: sum over the group's modes of (m, k, n) * R_p, [p][q] order
    const double Rs[4] = {R0, R1, R2, R3};
    const int wv = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const double vm = wave_sum(mm * Rs[p]);
        const double vk = wave_sum(kk * Rs[p]);
        const double vn = wave_sum(nn * Rs[p]);
        if (lane == 0) {
            s_red[wv][p * 3 + 0] = vm;
            s_red[wv][p * 3 + 1] = vk;
            s_red[wv][p * 3 + 2] = vn;
        }
    }
    __syncthreads();
    if (tid < 12) {
        double v = 0.0;
#pragma unroll
        for (int w = 0; w < ADJ_MODES / 64; w++) v += s_red[w][tid];
        a.part_phase[((long)c * a.groups + blockIdx.y) * 12 + tid] = v;
    }
}

// Interval j (blockIdx.y; j = L-1 and empty intervals give zeros) adds its
// chunks in order.
__global__ __launch_bounds__(256) void adjoint_reduce_kernel(AdjReduceArgs r)
{
    const int j = blockIdx.y;
    const long K2 = 2L * r.K;
    int c0 = 0, c1 = 0;
    if (j < r.nseg) {
        c0 = r.prefix[j];
        c1 = r.prefix[j + 1];
    }
    const long s = (long)blockIdx.x * 256 + threadIdx.x;
    if (s < K2) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int c = c0; c < c1; c++) {
            const double *pa = r.part_amp + (long)c * 4 * K2 + s;
#pragma unroll
            for (int p = 0; p < 4; p++) acc[p] += pa[p * K2];
        }
#pragma unroll
        for (int p = 0; p < 4; p++) r.g_interp[((long)p * r.L + j) * K2 + s] = acc[p];
    }
    if (blockIdx.x == 0 && threadIdx.x < 12) {
        const int v = threadIdx.x;
        double acc = 0.0;
        for (int c = c0; c < c1; c++)
            for (int g = 0; g < r.groups; g++) acc += r.part_phase[((long)c * r.groups + g) * 12 + v];
        const int p = v / 3, q = v % 3;
        r.g_phase[((long)p * r.L + j) * 3 + q] = acc;
    }
}

size_t adjoint_workspace_bytes(int L, int out_len, int K, int chunk, int groups)
{
    const long max_chunks = (long)(out_len + chunk - 1) / chunk + L;
    size_t b = 0;
    b += ((sizeof(Segment) * (size_t)L + 255) / 256) * 256;
    b += ((sizeof(int) * (size_t)(L + 1) + 255) / 256) * 256;
    b += ((sizeof(double) * (size_t)max_chunks * 8 * K + 255) / 256) * 256;
    b += ((sizeof(double) * (size_t)max_chunks * groups * 12 + 255) / 256) * 256;
    return b;
}

int launch_adjoint_fp64(const AdjArgs &a, int max_chunks, hipStream_t stream)
{
    if (max_chunks <= 0 || a.K <= 0) return FEW_OK;
    hipLaunchKernelGGL(adjoint_fp64_kernel, dim3(max_chunks, a.groups), dim3(ADJ_MODES), 0, stream,
                       a);
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

int launch_adjoint_reduce(const AdjReduceArgs &r, hipStream_t stream)
{
    const long K2 = 2L * r.K;
    const int bx = (int)((K2 + 255) / 256);
    hipLaunchKernelGGL(adjoint_reduce_kernel, dim3(bx > 0 ? bx : 1, r.L), dim3(256), 0, stream, r);
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

int launch_spline_adjoint(const SplineAdjArgs &a, hipStream_t stream)
{
    long total = 0;
    for (int g = 0; g < a.ngroups; g++) total += a.grp[g].nsys;
    if (total <= 0) return FEW_OK;
    if (total > a.ns) {
        set_error("spline_adjoint: %ld systems exceed %ld scratch columns", total, a.ns);
        return FEW_ERR_ARG;
    }
    if (a.L <= ADJ_SHARED_MAX_L) {
        SplineAdjArgs b = a;
        b.blk_prefix[0] = 0;
        for (int g = 0; g < 3; g++)
            b.blk_prefix[g + 1] =
                b.blk_prefix[g] + (g < a.ngroups ? (a.grp[g].nsys + ADJ_SPB - 1) / ADJ_SPB : 0);
        const size_t lds = sizeof(double) * (size_t)a.L * (8 + 2 * ADJ_SPB);
        FEW_LDS(spline_adjoint_lds_kernel, (int)lds);
        hipLaunchKernelGGL(spline_adjoint_lds_kernel, dim3((unsigned)b.blk_prefix[a.ngroups]),
                           dim3(256), lds, stream, b);
        FEW_HIP(hipGetLastError());
        return FEW_OK;
    }
    hipLaunchKernelGGL(tri_factor_kernel, dim3(1), dim3(64), 0, stream, a.t, a.L, a.factor,
                       nullptr);
    FEW_HIP(hipGetLastError());
    hipLaunchKernelGGL(spline_adjoint_kernel, dim3((unsigned)((total + 63) / 64)), dim3(64), 0,
                       stream, a);
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

}  // namespace few

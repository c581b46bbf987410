// Factorised (MFMA) interpolated mode sum for gfx950.
//
// See modesum_fact.hpp for the algebra.  Kernels:
//   fact_prep_kernel -- per knot interval, gathers C_p[r, n] (complex) from
//       the amplitude splines and Y_lm in a fixed order per cell (FP64
//       accumulation, scaled by h_j^p), and stores it as an FP32 tile laid
//       out exactly as the sum kernel keeps it in LDS.
//   fact_f32_kernel<RB, NB> -- one workgroup = 4 waves = one tile of 64*NB
//       samples inside one interval.  The interval's coefficient tile is
//       staged in LDS once; each wave owns NB 16-sample column blocks and
//       accumulates, with v_mfma_f32_16x16x4f32,
//           W[(r,p), t] = sum_n C_p[r, n] v_n(t)
//       as a real GEMM with the 2x2 [[Re, -Im], [Im, Re]] expansion.  B
//       fragments v_n(t) are produced in registers by an FP64 recurrence
//       v_{n+2} = v_n * exp(-2 i Phi_r) (no per-n transcendental), and the
//       epilogue forms h(t) = sum_{r,p} s^p u_r(t) W[(r,p), t] with the
//       products in FP32 and the sum in FP64, then adds into the complex128
//       output.  Phases (dense-output evaluation, the reductions to turns)
//       stay FP64.
//
// Precision class: "mixed" -- FP64 phases and spline build, FP32 amplitude
// coefficients / phasors / MFMA accumulation, FP64 output accumulation.

#include <math.h>

#include <algorithm>
#include <map>

#include "few_math.hpp"
#include "modesum_direct.hpp"
#include "modesum_fact.hpp"

namespace few {

typedef float v4f __attribute__((ext_vector_type(4)));

#define INV_2PI 0.15915494309189533577

// ---------------------------------------------------------------- host plan
bool build_fact_plan(const int *m, const int *k, const int *n, int K, FactPlanHost &p)
{
    p = FactPlanHost();
    p.K = K;
    if (K <= 0) return false;
    std::map<std::pair<int, int>, int> rows;
    int nlo = 1 << 30, nhi = -(1 << 30);
    for (int q = 0; q < K; q++) {
        rows[{m[q], k[q]}] = 0;
        nlo = std::min(nlo, n[q]);
        nhi = std::max(nhi, n[q]);
        if (m[q] != 0) {
            rows[{-m[q], -k[q]}] = 0;
            nlo = std::min(nlo, -n[q]);
            nhi = std::max(nhi, -n[q]);
        }
    }
    int r = 0;
    for (auto &kv : rows) {
        kv.second = r++;
        p.row_m.push_back(kv.first.first);
        p.row_k.push_back(kv.first.second);
    }
    p.Nr = r;
    p.n_lo = nlo;
    p.Nc = nhi - nlo + 1;
    p.RB = (8 * p.Nr + 15) / 16;
    p.tcols = p.Nc;
    p.tn0 = nlo;
    const int nmax = std::max(-nlo, nhi);
    if (nlo < 0 && nmax + 1 < p.Nc) {
        p.fold = 1;
        p.tcols = nmax + 1;
        p.tn0 = 0;
    }
    p.KS = (2 * p.tcols + 3) / 4;
    p.CST = (((p.fold ? 4 : 2) * p.KS + 15) / 16) * 16 + 2;
    const int ncell = p.Nr * p.Nc;
    std::vector<int> count(ncell + 1, 0);
    auto cell_of = [&](int mm, int kk, int nn) { return rows[{mm, kk}] * p.Nc + (nn - nlo); };
    for (int q = 0; q < K; q++) {
        count[cell_of(m[q], k[q], n[q]) + 1]++;
        if (m[q] != 0) count[cell_of(-m[q], -k[q], -n[q]) + 1]++;
    }
    p.cell_start.assign(ncell + 1, 0);
    for (int c = 0; c < ncell; c++) p.cell_start[c + 1] = p.cell_start[c] + count[c + 1];
    p.entries.assign(p.cell_start[ncell], 0);
    std::vector<int> fill(p.cell_start.begin(), p.cell_start.end() - 1);
    p.mode_cell.assign(2 * (size_t)K, -1);
    for (int q = 0; q < K; q++) {
        p.mode_cell[2 * q] = cell_of(m[q], k[q], n[q]);
        p.entries[fill[p.mode_cell[2 * q]]++] = 2 * q;
        if (m[q] != 0) {
            p.mode_cell[2 * q + 1] = cell_of(-m[q], -k[q], -n[q]);
            p.entries[fill[p.mode_cell[2 * q + 1]]++] = 2 * q + 1;
        }
    }
    p.rows_arith = 1;
    for (int rr = 0; rr < p.Nr; rr++)
        if (p.row_m[rr] != p.row_m[0] + rr || p.row_k[rr] != p.row_k[0]) p.rows_arith = 0;
    p.m0 = p.row_m[0];
    p.k0 = p.row_k[0];
    p.has_k = 0;
    for (int q = 0; q < K; q++)
        if (k[q] != 0) p.has_k = 1;
    p.ok = (p.RB <= FACT_MAX_RB) && (p.Nc <= FACT_MAX_NC);
    return p.ok;
}

// ---------------------------------------------------------------- prep kernel
__global__ __launch_bounds__(256) void fact_prep_kernel(FactPrepArgs a)
{
    const FactPlanDev P = a.plan;
    const long per_j = 2L * P.RB * P.CST;
    const long gid = (long)blockIdx.x * 256 + threadIdx.x;
    const int j = (int)(gid / per_j);
    if (j >= a.L - 1) return;
    const long rem = gid - (long)j * per_j;
    const int r = (int)(rem / P.CST);
    const int col = (int)(rem - (long)r * P.CST);
    // unfolded: pair col = (x0, x2) = (Re C, Im C); folded: pairs col and
    // TW + col = (x0, x2) and (x1, x3) of tile column col < TW, zero beyond 2 TW
    const int TW = 2 * P.KS;
    if (P.fold && col >= TW && col < 2 * TW) return;  // written by thread col - TW
    double x[4][4];
    fact_cell(P, a.interp, a.ylm, a.L, j, a.h_t[j + 1] - a.h_t[j], r, P.fold && col >= TW ? P.tcols : col,
              x);
    const bool second = P.fold && col < TW;
    if (a.C64) {
        double *dst = a.C64 + (long)j * a.c_stride;
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const long o = ((long)(r * 4 + p) * P.CST + col) * 2;
            dst[o] = x[p][0];
            dst[o + 1] = x[p][2];
            if (second) {
                dst[o + 2 * TW] = x[p][1];
                dst[o + 2 * TW + 1] = x[p][3];
            }
        }
        return;
    }
    float *dst = a.C + (long)j * a.c_stride;
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const long o = ((long)(r * 4 + p) * P.CST + col) * 2;
        dst[o] = (float)x[p][0];
        dst[o + 1] = (float)x[p][2];
        if (second) {
            dst[o + 2 * TW] = (float)x[p][1];
            dst[o + 2 * TW + 1] = (float)x[p][3];
        }
    }
}

int launch_fact_prep(const FactPrepArgs &a, hipStream_t stream)
{
    const long threads = (long)(a.L - 1) * 2 * a.plan.RB * a.plan.CST;
    if (threads <= 0) return FEW_OK;
    const long blocks = (threads + 255) / 256;
    hipLaunchKernelGGL(fact_prep_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, a);
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

// ---------------------------------------------------------------- sum kernel
__device__ __forceinline__ int fact_find_segment(const int *prefix, int nseg, int b)
{
    int lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
}

// One 64*NB-sample tile of interval j (coefficients already in LDS).
template <int RB, int NB>
__device__ __forceinline__ void fact_tile_body(const FactArgs &a, const FactPlanDev &P,
                                               const float *sC, const Segment &sg, int q,
                                               const double *sPh, int wv, int li, int g)
{
    // sPh: the interval's dense-output phase coefficients in LDS,
    // [0..7] Phi_phi, [8..15] Phi_theta, [16..23] Phi_r
    const double *rp = sPh;
    const double *rt = sPh + 8;
    const double *rr = sPh + 16;
    constexpr int TILE = 64 * NB;
    int idx[NB];
    bool ok[NB];
    double v_re[NB], v_im[NB], w_re[NB], w_im[NB], uph[NB], uth[NB];
    float sa[NB];
    const int n0 = P.tn0 + (g >> 1);
#pragma unroll
    for (int c = 0; c < NB; c++) {
        idx[c] = sg.start + q * TILE + wv * (16 * NB) + c * 16 + li;
        ok[c] = idx[c] < sg.end;
        const double t = a.dt * (double)idx[c];
        sa[c] = (float)((t - sg.h_t) / sg.hw);
        const double s = (t - sg.pt) / sg.pw;
        const double s1 = 1.0 - s;
        const double ur = dense_phase(rr, s, s1) * INV_2PI;  // Phi_r in turns
        uph[c] = dense_phase(rp, s, s1) * INV_2PI;
        uth[c] = dense_phase(rt, s, s1) * INV_2PI;
        double sn, cs;
        sincos_turns(reduce_turns((double)n0 * ur), sn, cs);
        v_re[c] = cs;   // v_n = exp(-i n Phi_r)
        v_im[c] = -sn;
        sincos_turns(reduce_turns(2.0 * ur), sn, cs);
        w_re[c] = cs;   // exp(-2 i Phi_r)
        w_im[c] = -sn;
    }

    v4f acc[RB][NB];
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
        for (int c = 0; c < NB; c++) acc[rb][c] = v4f{0.f, 0.f, 0.f, 0.f};

    const int e = li & 1;
    const int f = g & 1;
    // A entry (row 2R + e, column 2 acol + f): [[Re, -Im], [Im, Re]] of the
    // complex tile, or the folded tile's half f of the row
    const int pidx = P.fold ? f * 4 * P.KS + e : e ^ f;
    const float sgn = (!P.fold && e == 0 && f == 1) ? -1.f : 1.f;
    const int rbase = li >> 1;
    const int cst = P.CST;
    const int KS = P.KS;
    const int Nc = P.tcols;
    for (int s = 0; s < KS; s++) {
        const int acol = 2 * s + (g >> 1);
        float bv[NB];
#pragma unroll
        for (int c = 0; c < NB; c++) {
            const float val = f ? (float)v_im[c] : (float)v_re[c];
            bv[c] = (acol < Nc) ? val : 0.f;
            const double nr = v_re[c] * w_re[c] - v_im[c] * w_im[c];
            const double ni = v_re[c] * w_im[c] + v_im[c] * w_re[c];
            v_re[c] = nr;
            v_im[c] = ni;
        }
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            const int R = rb * 8 + rbase;
            const float av = sgn * sC[(R * cst + acol) * 2 + pidx];
#pragma unroll
            for (int c = 0; c < NB; c++)
                acc[rb][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv[c], acc[rb][c], 0, 0, 0);
        }
    }

    // epilogue: h(t) = sum_{r,p} s^p u_r(t) W[(r,p), t]; this lane holds rows
    // r = 2 rb + (g >> 1), powers p0 = 2 (g & 1) and p0 + 1
    const int p0 = 2 * f;
#pragma unroll
    for (int c = 0; c < NB; c++) {
        const float s1f = sa[c];
        const float s2f = s1f * s1f;
        const float spa = p0 == 0 ? 1.f : s2f;
        const float spb = p0 == 0 ? s1f : s2f * s1f;
        double hre = 0.0, him = 0.0;
        double u_re = 0.0, u_im = 0.0, e_re = 0.0, e_im = 0.0;
        if (P.rows_arith) {
            // u_r = exp(-2 pi i (m_r uph + k0 uth)), m_r = m0 + r: step by 2 rows
            double sn, cs;
            const double m_first = (double)(P.m0 + (g >> 1));
            sincos_turns(reduce_turns(m_first * uph[c] + (double)P.k0 * uth[c]), sn, cs);
            u_re = cs;
            u_im = -sn;
            sincos_turns(reduce_turns(2.0 * uph[c]), sn, cs);
            e_re = cs;
            e_im = -sn;
        }
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            const int r = 2 * rb + (g >> 1);
            float ur, ui;
            if (P.rows_arith) {
                ur = (float)u_re;
                ui = (float)u_im;
                const double nr = u_re * e_re - u_im * e_im;
                const double ni = u_re * e_im + u_im * e_re;
                u_re = nr;
                u_im = ni;
            } else {
                const int2 mk = r < P.Nr ? P.row_mk[r] : make_int2(0, 0);
                const double tu = reduce_turns((double)mk.x * uph[c] + (double)mk.y * uth[c]);
                float sn, cs;
                sincospif(2.0f * (float)tu, &sn, &cs);
                ur = cs;
                ui = -sn;
            }
            if (r < P.Nr) {
                const v4f W = acc[rb][c];
                const float zr = spa * W[0] + spb * W[2];
                const float zi = spa * W[1] + spb * W[3];
                hre += (double)(ur * zr - ui * zi);
                him += (double)(ur * zi + ui * zr);
            }
        }
        hre += __shfl_xor(hre, 16);
        him += __shfl_xor(him, 16);
        hre += __shfl_xor(hre, 32);
        him += __shfl_xor(him, 32);
        if (g == 0 && ok[c]) {
            double2 w = a.wave[idx[c]];
            w.x += hre;
            w.y += him;
            a.wave[idx[c]] = w;
        }
    }
}

// Persistent: block b owns the contiguous tile range [b*per, (b+1)*per) of
// the (interval, tile) sequence, so consecutive tiles share an interval and
// the LDS coefficient tile is reloaded only when the interval changes.
template <int RB, int NB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NB <= 2 ? 2 : 1, NB <= 2 ? 2 : 1)))
void fact_f32_kernel(FactArgs a)
{
    extern __shared__ float sC[];
    double *sPh = reinterpret_cast<double *>(sC + a.c_stride);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = tid >> 6;
    const int li = lane & 15;
    const int g = lane >> 4;
    const FactPlanDev P = a.plan;
    const int total = a.prefix[a.nseg];
    const int per = (total + (int)gridDim.x - 1) / (int)gridDim.x;
    const int t0 = (int)blockIdx.x * per;
    const int t1 = min(total, t0 + per);
    if (t0 >= t1) return;
    int j = fact_find_segment(a.prefix, a.nseg, t0);
    int jnext_start = a.prefix[j + 1];
    const long np = (long)a.nseg * 3;
    bool stale = true;
    for (int tile = t0; tile < t1; tile++) {
        while (tile >= jnext_start) {  // advance to the interval holding this tile
            j++;
            jnext_start = a.prefix[j + 1];
            stale = true;
        }
        if (stale) {
            __syncthreads();  // every wave is done reading the previous tile
            const float4 *src = reinterpret_cast<const float4 *>(a.C + (long)j * a.c_stride);
            float4 *dst = reinterpret_cast<float4 *>(sC);
            const int nv = (int)(a.c_stride >> 2);
            for (int i = tid; i < nv; i += 256) dst[i] = src[i];
            if (tid < 24) sPh[tid] = a.pcoef[(tid & 7) * np + (long)j * 3 + (tid >> 3)];
            __syncthreads();
            stale = false;
        }
        const Segment sg = a.segs[j];
        fact_tile_body<RB, NB>(a, P, sC, sg, tile - a.prefix[j], sPh, wv, li, g);
    }
}

// ---------------------------------------------------------------- FP64 sum
// fact_tile_body in FP64: the tile entries, the v_n and u_r recurrences and
// the epilogue stay FP64, and the n-contraction runs on
// v_mfma_f64_16x16x4f64 (FP64 products and accumulation).  That MFMA returns
// D[g + 4 i][l & 15] to lane l (g = l >> 4) where the FP32 one returns
// D[4 g + i]; the A rows are permuted so that a lane still receives
// (p0 Re, p0 Im, p0+1 Re, p0+1 Im) of row r = 2 rb + (g >> 1): MFMA row rho
// carries complex row r_local = (rho & 3) >> 1, power 2 (rho & 1) + (rho >> 3)
// and part (rho >> 2) & 1.
typedef double v4d __attribute__((ext_vector_type(4)));

template <int RB, int NB, int WAVES>
__device__ __forceinline__ void fact64_tile_body(const FactArgs &a, const FactPlanDev &P,
                                                 const double *sC, const Segment &sg, int q,
                                                 const double *sPh, int wv, int li, int g)
{
    // rounding fixed by the source, not by the compiler's fusion choices, so
    // every variant (samples per wave, waves) gives the same bits
#pragma clang fp contract(off)
    const double *rp = sPh;
    const double *rt = sPh + 8;
    const double *rr = sPh + 16;
    constexpr int TILE = 16 * NB * WAVES;
    int idx[NB];
    bool ok[NB];
    double v_re[NB], v_im[NB], w_re[NB], w_im[NB], uph[NB], uth[NB], sa[NB];
    const int n0 = P.tn0 + (g >> 1);
#pragma unroll
    for (int c = 0; c < NB; c++) {
        idx[c] = sg.start + q * TILE + wv * (16 * NB) + c * 16 + li;
        ok[c] = idx[c] < sg.end;
        const double t = a.dt * (double)idx[c];
        sa[c] = (t - sg.h_t) / sg.hw;
        const double s = (t - sg.pt) / sg.pw;
        const double s1 = 1.0 - s;
        const double ur = dense_phase(rr, s, s1) * INV_2PI;
        uph[c] = dense_phase(rp, s, s1) * INV_2PI;
        uth[c] = P.has_k ? dense_phase(rt, s, s1) * INV_2PI : 0.0;
        double sn, cs;
        sincos_turns(reduce_turns((double)n0 * ur), sn, cs);
        v_re[c] = cs;
        v_im[c] = -sn;
        sincos_turns(reduce_turns(2.0 * ur), sn, cs);
        w_re[c] = cs;
        w_im[c] = -sn;
    }
    v4d acc[RB][NB];
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
        for (int c = 0; c < NB; c++) acc[rb][c] = v4d{0.0, 0.0, 0.0, 0.0};
    const int f = g & 1;
    const int r_loc = (li & 3) >> 1, p_row = 2 * (li & 1) + (li >> 3), e = (li >> 2) & 1;
    const int pidx = P.fold ? f * 4 * P.KS + e : e ^ f;
    const double sgn = (!P.fold && e == 0 && f == 1) ? -1.0 : 1.0;
    const int cst = P.CST;
    const int KS = P.KS;
    const int Nc = P.tcols;
    for (int s = 0; s < KS; s++) {
        const int acol = 2 * s + (g >> 1);
        double bv[NB];
#pragma unroll
        for (int c = 0; c < NB; c++) {
            bv[c] = (acol < Nc) ? (f ? v_im[c] : v_re[c]) : 0.0;
            const double nr = v_re[c] * w_re[c] - v_im[c] * w_im[c];
            const double ni = v_re[c] * w_im[c] + v_im[c] * w_re[c];
            v_re[c] = nr;
            v_im[c] = ni;
        }
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            const int R = (2 * rb + r_loc) * 4 + p_row;
            const double av = sgn * sC[(R * cst + acol) * 2 + pidx];
#pragma unroll
            for (int c = 0; c < NB; c++)
                acc[rb][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[c], acc[rb][c], 0, 0, 0);
        }
    }
    const int p0 = 2 * f;
#pragma unroll
    for (int c = 0; c < NB; c++) {
        const double s1 = sa[c], s2 = s1 * s1;
        const double spa = p0 == 0 ? 1.0 : s2;
        const double spb = p0 == 0 ? s1 : s2 * s1;
        double hre = 0.0, him = 0.0;
        double u_re = 0.0, u_im = 0.0, e_re = 0.0, e_im = 0.0;
        if (P.rows_arith) {
            double sn, cs;
            const double m_first = (double)(P.m0 + (g >> 1));
            sincos_turns(reduce_turns(m_first * uph[c] + (double)P.k0 * uth[c]), sn, cs);
            u_re = cs;
            u_im = -sn;
            sincos_turns(reduce_turns(2.0 * uph[c]), sn, cs);
            e_re = cs;
            e_im = -sn;
        }
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            const int r = 2 * rb + (g >> 1);
            double ur, ui;
            if (P.rows_arith) {
                ur = u_re;
                ui = u_im;
                const double nr = u_re * e_re - u_im * e_im;
                const double ni = u_re * e_im + u_im * e_re;
                u_re = nr;
                u_im = ni;
            } else {
                const int2 mk = r < P.Nr ? P.row_mk[r] : make_int2(0, 0);
                double sn, cs;
                sincos_turns(reduce_turns((double)mk.x * uph[c] + (double)mk.y * uth[c]), sn, cs);
                ur = cs;
                ui = -sn;
            }
            if (r < P.Nr) {
                const v4d W = acc[rb][c];
                const double zr = spa * W[0] + spb * W[2];
                const double zi = spa * W[1] + spb * W[3];
                hre += ur * zr - ui * zi;
                him += ur * zi + ui * zr;
            }
        }
        hre += __shfl_xor(hre, 16);
        him += __shfl_xor(him, 16);
        hre += __shfl_xor(hre, 32);
        him += __shfl_xor(him, 32);
        if (g == 0 && ok[c]) {
            if (a.overwrite) {
                a.wave[idx[c]] = make_double2(hre, him);
            } else {
                double2 w = a.wave[idx[c]];
                w.x += hre;
                w.y += him;
                a.wave[idx[c]] = w;
            }
        }
    }
}

template <int RB, int NB, int WAVES>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(WAVES / 4, WAVES / 4)))
void fact_f64_kernel(FactArgs a)
{
    constexpr int NT = 64 * WAVES;
    extern __shared__ __attribute__((aligned(16))) double sC64[];
    double *sPh = sC64 + a.c_stride;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = tid >> 6;
    const int li = lane & 15;
    const int g = lane >> 4;
    const FactPlanDev P = a.plan;
    const int total = a.prefix[a.nseg];
    const int per = (total + (int)gridDim.x - 1) / (int)gridDim.x;
    const int t0 = (int)blockIdx.x * per;
    const int t1 = min(total, t0 + per);
    if (a.overwrite) {
        // samples after the last interval: zero (the intervals cover [0, end))
        const long end = a.nseg > 0 ? a.segs[a.nseg - 1].end : 0;
        for (long i = end + (long)blockIdx.x * NT + tid; i < a.out_len; i += (long)gridDim.x * NT)
            a.wave[i] = make_double2(0.0, 0.0);
    }
    if (t0 >= t1) return;
    int j = fact_find_segment(a.prefix, a.nseg, t0);
    int jnext_start = a.prefix[j + 1];
    const long np = (long)a.nseg * 3;
    bool stale = true;
    for (int tile = t0; tile < t1; tile++) {
        while (tile >= jnext_start) {
            j++;
            jnext_start = a.prefix[j + 1];
            stale = true;
        }
        if (stale) {
            __syncthreads();
            const double2 *src = reinterpret_cast<const double2 *>(a.C64 + (long)j * a.c_stride);
            double2 *dst = reinterpret_cast<double2 *>(sC64);
            const int nv = (int)(a.c_stride >> 1);
            for (int i = tid; i < nv; i += NT) dst[i] = src[i];
            if (tid < 24) sPh[tid] = a.pcoef[(tid & 7) * np + (long)j * 3 + (tid >> 3)];
            __syncthreads();
            stale = false;
        }
        const Segment sg = a.segs[j];
        fact64_tile_body<RB, NB, WAVES>(a, P, sC64, sg, tile - a.prefix[j], sPh, wv, li, g);
    }
}

// variant 0: 8 waves x 16 samples (two waves per SIMD); 4: 4 waves x 32
// samples (one wave per SIMD with the whole register file)
int fact64_tile(int /*variant*/) { return 128; }

int fact_tile(const FactPlanHost & /*p*/, int variant) { return 64 * (variant == 4 ? 4 : 2); }

static int device_cus()
{
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cached[dev]) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        cached[dev] = n;
    }
    return cached[dev];
}

template <int NB>
static int launch_fact_nb(const FactArgs &a, int RB, int blocks, size_t lds, hipStream_t stream)
{
    switch (RB) {
#define FEW_CASE(X)                                                                            \
    case X:                                                                                    \
        if (lds > 65536)                                                                       \
            FEW_LDS((fact_f32_kernel<X, NB>), lds);                                            \
        hipLaunchKernelGGL((fact_f32_kernel<X, NB>), dim3(blocks), dim3(256), lds, stream, a); \
        break;
        FEW_CASE(1) FEW_CASE(2) FEW_CASE(3) FEW_CASE(4) FEW_CASE(5) FEW_CASE(6) FEW_CASE(7)
        FEW_CASE(8) FEW_CASE(9) FEW_CASE(10) FEW_CASE(11) FEW_CASE(12) FEW_CASE(13) FEW_CASE(14)
        FEW_CASE(15) FEW_CASE(16)
#undef FEW_CASE
    default:
        set_error("factorised sum: %d row blocks unsupported", RB);
        return FEW_ERR_UNSUPPORTED;
    }
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

template <int NB, int WAVES>
static int launch_fact64_nb(const FactArgs &a, int RB, int blocks, size_t lds, hipStream_t stream)
{
    switch (RB) {
#define FEW_CASE64(X)                                                                          \
    case X:                                                                                    \
        FEW_LDS((fact_f64_kernel<X, NB, WAVES>), lds);                                         \
        hipLaunchKernelGGL((fact_f64_kernel<X, NB, WAVES>), dim3(blocks), dim3(64 * WAVES), lds, \
                           stream, a);                                                         \
        break;
        FEW_CASE64(1) FEW_CASE64(2) FEW_CASE64(3) FEW_CASE64(4) FEW_CASE64(5) FEW_CASE64(6)
        FEW_CASE64(7) FEW_CASE64(8) FEW_CASE64(9) FEW_CASE64(10) FEW_CASE64(11) FEW_CASE64(12)
        FEW_CASE64(13) FEW_CASE64(14) FEW_CASE64(15) FEW_CASE64(16)
#undef FEW_CASE64
    default:
        set_error("fact64 sum: %d row blocks unsupported", RB);
        return FEW_ERR_UNSUPPORTED;
    }
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

int launch_fact64_sum(const FactArgs &a, const FactPlanHost &p, int max_blocks, int variant,
                      hipStream_t stream)
{
    if (max_blocks <= 0 && !a.overwrite) return FEW_OK;
    if (variant != 0 && variant != 4) {
        set_error("fact64 sum: unknown variant %d", variant);
        return FEW_ERR_ARG;
    }
    const size_t lds = (size_t)fact_coef_floats_per_interval(p) * sizeof(double) + 24 * sizeof(double);
    if (lds > 160 * 1024) {
        set_error("fact64 sum: coefficient tile of %zu bytes exceeds LDS", lds);
        return FEW_ERR_UNSUPPORTED;
    }
    int blocks = device_cus() * (lds <= 80 * 1024 ? 2 : 1);
    if (blocks > max_blocks) blocks = max_blocks;
    if (blocks < 1) blocks = 1;
    if (variant == 4) return launch_fact64_nb<2, 4>(a, p.RB, blocks, lds, stream);
    return launch_fact64_nb<1, 8>(a, p.RB, blocks, lds, stream);
}

int launch_fact_sum(const FactArgs &a, const FactPlanHost &p, int max_blocks, int variant,
                    hipStream_t stream)
{
    if (max_blocks <= 0) return FEW_OK;
    const size_t lds = (size_t)fact_coef_floats_per_interval(p) * sizeof(float) + 24 * sizeof(double);
    if (lds > 160 * 1024) {
        set_error("factorised sum: coefficient tile of %zu bytes exceeds LDS", lds);
        return FEW_ERR_UNSUPPORTED;
    }
    // persistent grid: enough resident blocks to fill every CU
    const int per_cu = (variant == 4) ? 1 : (lds <= 80 * 1024 ? 2 : 1);
    int blocks = device_cus() * per_cu;
    if (blocks > max_blocks) blocks = max_blocks;
    if (variant == 4) return launch_fact_nb<4>(a, p.RB, blocks, lds, stream);
    return launch_fact_nb<2>(a, p.RB, blocks, lds, stream);
}

}  // namespace few

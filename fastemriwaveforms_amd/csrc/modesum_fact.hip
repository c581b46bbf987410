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
    p.KS = (2 * p.Nc + 3) / 4;
    p.CST = ((2 * p.KS + 15) / 16) * 16 + 2;
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
    for (int q = 0; q < K; q++) {
        p.entries[fill[cell_of(m[q], k[q], n[q])]++] = 2 * q;
        if (m[q] != 0) p.entries[fill[cell_of(-m[q], -k[q], -n[q])]++] = 2 * q + 1;
    }
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
    double cre[4] = {0.0, 0.0, 0.0, 0.0}, cim[4] = {0.0, 0.0, 0.0, 0.0};
    if (r < P.Nr && col < P.Nc) {
        const int cell = r * P.Nc + col;
        const long K2 = 2L * P.K;
        for (int e = P.cell_start[cell]; e < P.cell_start[cell + 1]; e++) {
            const int ent = P.entries[e];
            const int qm = ent >> 1;
            const bool cj = ent & 1;
            const double2 Y = a.ylm[cj ? P.K + qm : qm];
#pragma unroll
            for (int p = 0; p < 4; p++) {
                const long base = ((long)p * a.L + j) * K2;
                const double cr = a.interp[base + qm];
                const double ci = cj ? -a.interp[base + P.K + qm] : a.interp[base + P.K + qm];
                cre[p] += cr * Y.x - ci * Y.y;
                cim[p] += cr * Y.y + ci * Y.x;
            }
        }
        const double hw = a.h_t[j + 1] - a.h_t[j];
        double hp = hw;
#pragma unroll
        for (int p = 1; p < 4; p++) {
            cre[p] *= hp;
            cim[p] *= hp;
            hp *= hw;
        }
    }
    float *dst = a.C + (long)j * a.c_stride;
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const long o = ((long)(r * 4 + p) * P.CST + col) * 2;
        dst[o] = (float)cre[p];
        dst[o + 1] = (float)cim[p];
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

template <int RB, int NB>
__global__ __launch_bounds__(256) void fact_f32_kernel(FactArgs a)
{
    extern __shared__ float sC[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = tid >> 6;
    const int li = lane & 15;
    const int g = lane >> 4;
    const int b = blockIdx.x;
    const int total = a.prefix[a.nseg];
    if (b >= total) return;
    const int j = fact_find_segment(a.prefix, a.nseg, b);
    const Segment sg = a.segs[j];
    const int q = b - a.prefix[j];
    const FactPlanDev P = a.plan;

    {   // stage the interval's coefficient tile
        const float4 *src = reinterpret_cast<const float4 *>(a.C + (long)j * a.c_stride);
        float4 *dst = reinterpret_cast<float4 *>(sC);
        const int nv = (int)(a.c_stride >> 2);
        for (int i = tid; i < nv; i += 256) dst[i] = src[i];
    }

    const long np = (long)a.nseg * 3;
    double rp[8], rt[8], rr[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        rp[c] = a.pcoef[c * np + (long)j * 3 + 0];
        rt[c] = a.pcoef[c * np + (long)j * 3 + 1];
        rr[c] = a.pcoef[c * np + (long)j * 3 + 2];
    }

    constexpr int TILE = 64 * NB;
    int idx[NB];
    bool ok[NB];
    double v_re[NB], v_im[NB], w_re[NB], w_im[NB], uph[NB], uth[NB];
    float sa[NB];
    const int n0 = P.n_lo + (g >> 1);
#pragma unroll
    for (int c = 0; c < NB; c++) {
        idx[c] = sg.start + q * TILE + wv * (16 * NB) + c * 16 + li;
        ok[c] = idx[c] < sg.end;
        const double t = a.dt * (double)idx[c];
        sa[c] = (float)((t - sg.h_t) / sg.hw);
        const double s = (t - sg.pt) / sg.pw;
        const double s1 = 1.0 - s;
        const double Pp = dense_phase(rp, s, s1);
        const double Pt = dense_phase(rt, s, s1);
        const double Pr = dense_phase(rr, s, s1);
        uph[c] = Pp * INV_2PI;
        uth[c] = Pt * INV_2PI;
        double sn, cs;
        sincos((double)n0 * Pr, &sn, &cs);
        v_re[c] = cs;
        v_im[c] = -sn;
        sincos(2.0 * Pr, &sn, &cs);
        w_re[c] = cs;
        w_im[c] = -sn;
    }
    __syncthreads();

    v4f acc[RB][NB];
#pragma unroll
    for (int rb = 0; rb < RB; rb++)
#pragma unroll
        for (int c = 0; c < NB; c++) acc[rb][c] = v4f{0.f, 0.f, 0.f, 0.f};

    const int e = li & 1;
    const int f = g & 1;
    const int pidx = e ^ f;
    const float sgn = (e == 0 && f == 1) ? -1.f : 1.f;
    const int rbase = li >> 1;
    const int cst = P.CST;
    const int KS = P.KS;
    const int Nc = P.Nc;
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

    // epilogue: h(t) = sum_{r,p} s^p u_r(t) W[(r,p), t]
    const int p0 = 2 * f;
#pragma unroll
    for (int c = 0; c < NB; c++) {
        const float s1f = sa[c];
        const float s2f = s1f * s1f;
        const float spa = p0 == 0 ? 1.f : s2f;
        const float spb = p0 == 0 ? s1f : s2f * s1f;
        double hre = 0.0, him = 0.0;
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
            const int r = 2 * rb + (g >> 1);
            if (r < P.Nr) {
                const v4f W = acc[rb][c];
                const float zr = spa * W[0] + spb * W[2];
                const float zi = spa * W[1] + spb * W[3];
                const int2 mk = P.row_mk[r];
                double tu = (double)mk.x * uph[c] + (double)mk.y * uth[c];
                tu -= rint(tu);
                float sn, cs;
                sincospif(2.0f * (float)tu, &sn, &cs);
                // u = cs - i sn
                hre += (double)(cs * zr + sn * zi);
                him += (double)(cs * zi - sn * zr);
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

int fact_tile(const FactPlanHost & /*p*/, int variant) { return 64 * (variant == 2 ? 2 : 4); }

template <int NB>
static int launch_fact_nb(const FactArgs &a, int RB, int blocks, size_t lds, hipStream_t stream)
{
    switch (RB) {
#define FEW_CASE(X)                                                                            \
    case X:                                                                                    \
        if (lds > 65536)                                                                       \
            FEW_HIP(hipFuncSetAttribute((const void *)fact_f32_kernel<X, NB>,                  \
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
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

int launch_fact_sum(const FactArgs &a, const FactPlanHost &p, int max_blocks, int variant,
                    hipStream_t stream)
{
    if (max_blocks <= 0) return FEW_OK;
    const size_t lds = (size_t)fact_coef_floats_per_interval(p) * sizeof(float);
    if (lds > 160 * 1024) {
        set_error("factorised sum: coefficient tile of %zu bytes exceeds LDS", lds);
        return FEW_ERR_UNSUPPORTED;
    }
    if (variant == 2) return launch_fact_nb<2>(a, p.RB, max_blocks, lds, stream);
    return launch_fact_nb<4>(a, p.RB, max_blocks, lds, stream);
}

}  // namespace few

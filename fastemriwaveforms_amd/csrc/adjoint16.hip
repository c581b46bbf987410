// Mixed-precision adjoint of the factorised mode sum ("mixed16" backward).
//
// With u_r = exp(-i(m Phi_phi + k Phi_theta)), v_n = exp(-i n Phi_r) (the
// factorisation of the forward) and sigma = (t - t_j)/h_j on knot interval j,
// every quantity the backward needs is a moment
//     M_j[r, n, s] = sum_{i in j} sigma_i^s conj(g_i) u_r(t_i) v_n(t_i),  s = 0..6,
// because (adjoint.hpp) dL/dA_q and dL/dphi_q are linear in conj(g) e_q and
// e_q = u_r v_n (its conjugate term: u_{-r} v_{-n}).  Per interval this is a
// (rows (r, s)) x (samples) x (samples) x (columns n) contraction -- the
// transposed twin of the forward GEMM -- run here on
// v_mfma_f32_16x16x32_f16 with both operands split into FP16 hi + lo parts
// (three products, FP32 accumulation), exactly like the forward kernel.
//
// One workgroup owns one chunk of <= A16_CHUNK samples of one interval and
// NBC waves (one 16-column block of n each).  Per 16-sample step the workgroup
// builds in LDS: the phasor tables u (Nr x 16), w = sigma^s * scale * conj(g)
// (7 x 16), the split v table (NBC*16 columns x 16 samples), and from u and w
// the split A tile (rows 2R, 2R+1 of complex row R = 7 r + s); then every
// wave runs its RBM x 3 MFMAs.  Phases are FP64 (turns, reduced), sincos is
// the FP64 sincos_turns.  The chunk's FP32 moments are unscaled (powers of
// two, exact) and written per chunk; later kernels add the chunks of each
// interval in FP64 in a fixed order and finish per mode:
//   dL/dRe c_p = h^p Re(Y+ M+_p + Y- M-_p),  dL/dIm c_p = h^p Im(Y- M-_p - Y+ M+_p),
//   sum_i sigma^p dL/dphi_q = Im(Y+ sum_s a_s M+_{p+s}) - Im(Y- sum_s conj(a_s) M-_{p+s}),
// a_s = c_s h^s, M+ at the mode's cell and M- at its conjugate cell.
#include <math.h>

#include "adjoint.hpp"
#include "batch.hpp"
#include "few_math.hpp"
#include "modesum_direct.hpp"
#include "modesum_fact.hpp"

namespace few {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

#define A16_SUPER 256          // samples whose phases are staged at a time
#define A16_AST 40             // A-tile row stride in halfs (32 + 8 padding)
#define A16_VST 40             // v-table column stride in halfs (32 + 8 padding: the 16 lanes of a
                               // fragment read start on distinct 4-bank groups)
#define INV_2PI_A 0.15915494309189533577

struct A16Layout {
    size_t a_hi, a_lo, v_hi, v_lo, u, w, pp, pt, pr, sig, cg, ph, red, total;
};

static __host__ __device__ inline size_t a16_align(size_t x) { return (x + 15) & ~(size_t)15; }

__host__ __device__ inline A16Layout a16_layout(int RBM, int NBC, int Nr)
{
    A16Layout L;
    size_t o = 0;
    L.a_hi = o; o = a16_align(o + (size_t)RBM * 16 * A16_AST * 2);
    L.a_lo = o; o = a16_align(o + (size_t)RBM * 16 * A16_AST * 2);
    L.v_hi = o; o = a16_align(o + (size_t)NBC * 16 * A16_VST * 2);
    L.v_lo = o; o = a16_align(o + (size_t)NBC * 16 * A16_VST * 2);
    L.u = o;    o = a16_align(o + (size_t)Nr * 16 * 8);
    L.w = o;    o = a16_align(o + (size_t)A16_MOM * 16 * 8);
    L.pp = o;   o = a16_align(o + (size_t)A16_SUPER * 8);
    L.pt = o;   o = a16_align(o + (size_t)A16_SUPER * 8);
    L.pr = o;   o = a16_align(o + (size_t)A16_SUPER * 8);
    L.sig = o;  o = a16_align(o + (size_t)A16_SUPER * 4);
    L.cg = o;   o = a16_align(o + (size_t)A16_SUPER * 8);
    L.ph = o;   o = a16_align(o + 24 * 8);
    L.red = o;  o = a16_align(o + 16 * 4);
    L.total = o;
    return L;
}

size_t adjoint16_lds_bytes(int RBM, int NBC, int Nr) { return a16_layout(RBM, NBC, Nr).total; }

__device__ __forceinline__ void split16(float x, _Float16 &h, _Float16 &l)
{
    h = (_Float16)x;
    l = (_Float16)(x - (float)h);
}

template <int RBM>
__global__ __launch_bounds__(512) void adjoint16_kernel(Adj16Args a)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const FactPlanDev &P = a.plan;
    const int NBC = a.NBC;
    const int NT = 64 * NBC;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6, li = lane & 15, g = lane >> 4;
    const int c = blockIdx.x;
    const int total = a.prefix[a.nseg];
    if (c >= total) return;  // uniform over the workgroup
    const int j = upper_index(a.prefix, a.nseg, c);
    const Segment sg = a.segs[j];
    const int lo = sg.start + (c - a.prefix[j]) * a.chunk;
    const int hi = min(lo + a.chunk, sg.end);
    const int Nr = P.Nr;
    const A16Layout Lay = a16_layout(RBM, NBC, Nr);
    _Float16 *sAh = reinterpret_cast<_Float16 *>(smem + Lay.a_hi);
    _Float16 *sAl = reinterpret_cast<_Float16 *>(smem + Lay.a_lo);
    _Float16 *sVh = reinterpret_cast<_Float16 *>(smem + Lay.v_hi);
    _Float16 *sVl = reinterpret_cast<_Float16 *>(smem + Lay.v_lo);
    float2 *sU = reinterpret_cast<float2 *>(smem + Lay.u);
    float2 *sW = reinterpret_cast<float2 *>(smem + Lay.w);
    double *sPp = reinterpret_cast<double *>(smem + Lay.pp);
    double *sPt = reinterpret_cast<double *>(smem + Lay.pt);
    double *sPr = reinterpret_cast<double *>(smem + Lay.pr);
    float *sSig = reinterpret_cast<float *>(smem + Lay.sig);
    float2 *sCg = reinterpret_cast<float2 *>(smem + Lay.cg);
    double *sPh = reinterpret_cast<double *>(smem + Lay.ph);
    float *sRed = reinterpret_cast<float *>(smem + Lay.red);

    // phase coefficients of interval j; chunk scale from max |g|
    const long np = (long)a.nseg * 3;
    if (tid < 24) sPh[tid] = a.pcoef[(tid & 7) * np + (long)j * 3 + (tid >> 3)];
    float mx = 0.f;
    for (int i = lo + tid; i < hi; i += NT) {
        const double2 gg = a.grad[i];
        mx = fmaxf(mx, fmaxf(fabsf((float)gg.x), fabsf((float)gg.y)));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    if (lane == 0) sRed[wv] = mx;
    // zero the padding rows of the A tile (complex rows >= 7 Nr) once
    const int rows_used = 2 * A16_MOM * Nr;
    for (int e = tid; e < (RBM * 16 - rows_used) * A16_AST; e += NT) {
        sAh[rows_used * A16_AST + e] = (_Float16)0.f;
        sAl[rows_used * A16_AST + e] = (_Float16)0.f;
    }
    __syncthreads();
    float gmax = 0.f;
    for (int w = 0; w < NBC; w++) gmax = fmaxf(gmax, sRed[w]);
    int ex = 0;
    if (gmax > 0.f) frexpf(gmax, &ex);  // gmax in [2^(ex-1), 2^ex)
    const float scale = ldexpf(1.f, 13 - ex);
    const float inv_out = ldexpf(1.f, ex - 13 - 8);  // 1 / (scale * 256)

    const double inv_pw = 1.0 / sg.pw;
    const double inv_hw = 1.0 / sg.hw;
    v4f acc[RBM];
#pragma unroll
    for (int rb = 0; rb < RBM; rb++) acc[rb] = v4f{0.f, 0.f, 0.f, 0.f};

    for (int sbase = lo; sbase < hi; sbase += A16_SUPER) {
        __syncthreads();
        for (int q = tid; q < A16_SUPER; q += NT) {
            const int i = sbase + q;
            const double t = a.dt * (double)i;
            const double s = (t - sg.pt) * inv_pw;
            sPp[q] = dense_phase(sPh, s, 1.0 - s) * INV_2PI_A;
            sPt[q] = P.has_k ? dense_phase(sPh + 8, s, 1.0 - s) * INV_2PI_A : 0.0;
            sPr[q] = dense_phase(sPh + 16, s, 1.0 - s) * INV_2PI_A;
            sSig[q] = (float)((t - sg.h_t) * inv_hw);
            float2 cg = make_float2(0.f, 0.f);
            if (i < hi) {
                const double2 gg = a.grad[i];
                cg = make_float2(scale * (float)gg.x, -scale * (float)gg.y);
            }
            sCg[q] = cg;
        }
        const int nsteps = (min(A16_SUPER, hi - sbase) + 15) >> 4;
        for (int st = 0; st < nsteps; st++) {
            const int i0 = st * 16;
            __syncthreads();
            // u_r(t_i)
            for (int e = tid; e < Nr * 16; e += NT) {
                const int r = e >> 4, ii = e & 15;
                const int2 mk = P.row_mk[r];
                double sn, cs;
                sincos_turns(reduce_turns((double)mk.x * sPp[i0 + ii] + (double)mk.y * sPt[i0 + ii]),
                             sn, cs);
                sU[e] = make_float2((float)cs, (float)(-sn));
            }
            // w_s = sigma^s * scale * conj(g)
            for (int e = tid; e < A16_MOM * 16; e += NT) {
                const int s = e >> 4, ii = e & 15;
                const float sg1 = sSig[i0 + ii];
                float pw = 1.f;
                for (int k2 = 0; k2 < s; k2++) pw *= sg1;
                const float2 cg = sCg[i0 + ii];
                sW[e] = make_float2(pw * cg.x, pw * cg.y);
            }
            // v_n(t_i) * 256, split: four consecutive columns per thread
            for (int e = tid; e < NBC * 4 * 16; e += NT) {
                const int ii = e & 15, cg4 = e >> 4;
                const double ph = sPr[i0 + ii];
                double sn, cs, zs, zc;
                sincos_turns(reduce_turns((double)(P.n_lo + 4 * cg4) * ph), sn, cs);
                sincos_turns(reduce_turns(ph), zs, zc);
                double vr = 256.0 * cs, vi = -256.0 * sn;
                const double zr = zc, zi = -zs;
#pragma unroll
                for (int t2 = 0; t2 < 4; t2++) {
                    const int col = 4 * cg4 + t2;
                    const bool in = col < P.Nc;
                    _Float16 hr, lr, hi_, li_;
                    split16(in ? f32_opaque((float)vr) : 0.f, hr, lr);
                    split16(in ? f32_opaque((float)vi) : 0.f, hi_, li_);
                    const int o = col * A16_VST + 2 * ii;
                    typedef _Float16 h2v __attribute__((ext_vector_type(2)));
                    *reinterpret_cast<h2v *>(sVh + o) = h2v{hr, hi_};
                    *reinterpret_cast<h2v *>(sVl + o) = h2v{lr, li_};
                    const double nr = vr * zr - vi * zi;
                    const double ni = vr * zi + vi * zr;
                    vr = nr;
                    vi = ni;
                }
            }
            __syncthreads();
            // A tile: complex row R = 7 r + s, sample ii: a = w_s u_r.  A thread
            // keeps its sample ii and walks the rows; each (re, im) pair of a
            // row is one 32-bit LDS store.
            {
                const int ii = tid & 15;
                for (int R = tid >> 4; R < A16_MOM * Nr; R += NT >> 4) {
                    const int r = R / A16_MOM, s = R - A16_MOM * r;
                    const float2 w = sW[s * 16 + ii];
                    const float2 u = sU[r * 16 + ii];
                    const float ar = w.x * u.x - w.y * u.y;
                    const float ai = w.x * u.y + w.y * u.x;
                    _Float16 hr, lr, hi_, li_;
                    split16(ar, hr, lr);
                    split16(ai, hi_, li_);
                    const int o0 = (2 * R) * A16_AST + 2 * ii;
                    const int o1 = o0 + A16_AST;
                    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
                    *reinterpret_cast<h2 *>(sAh + o0) = h2{hr, -hi_};
                    *reinterpret_cast<h2 *>(sAh + o1) = h2{hi_, hr};
                    *reinterpret_cast<h2 *>(sAl + o0) = h2{lr, -li_};
                    *reinterpret_cast<h2 *>(sAl + o1) = h2{li_, lr};
                }
            }
            __syncthreads();
            const v8h bh = *reinterpret_cast<const v8h *>(sVh + (16 * wv + li) * A16_VST + 8 * g);
            const v8h bl = *reinterpret_cast<const v8h *>(sVl + (16 * wv + li) * A16_VST + 8 * g);
            const _Float16 *aH = sAh + li * A16_AST + 8 * g;
            const _Float16 *aL = sAl + li * A16_AST + 8 * g;
#pragma unroll
            for (int rb = 0; rb < RBM; rb++) {
                const v8h ah = *reinterpret_cast<const v8h *>(aH + rb * 16 * A16_AST);
                const v8h al = *reinterpret_cast<const v8h *>(aL + rb * 16 * A16_AST);
                acc[rb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, acc[rb], 0, 0, 0);
                acc[rb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, acc[rb], 0, 0, 0);
                acc[rb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc[rb], 0, 0, 0);
            }
        }
    }
    // chunk moments: D[16 rb + 4 g + v][16 wv + li] / (scale * 256)
    const int ncol = NBC * 16;
    float *out = a.part + (long)c * RBM * 16 * ncol;
#pragma unroll
    for (int rb = 0; rb < RBM; rb++)
#pragma unroll
        for (int v = 0; v < 4; v++)
            out[(long)(16 * rb + 4 * g + v) * ncol + 16 * wv + li] = acc[rb][v] * inv_out;
}

// M_j = sum over the chunks of interval j (FP64, chunk order); zero for
// intervals without samples.
__global__ __launch_bounds__(256) void adjoint16_reduce_kernel(const float *part, const int *prefix,
                                                               int nseg, long tile, double *M)
{
    const int j = blockIdx.y;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= tile) return;
    double acc = 0.0;
    for (int c = prefix[j]; c < prefix[j + 1]; c++) acc += (double)part[(long)c * tile + e];
    M[(long)j * tile + e] = acc;
}

// Per (interval j, mode q): amplitude-coefficient gradients and the phase
// moments, reduced over the workgroup's modes.
__global__ __launch_bounds__(256) void adjoint16_finish_kernel(Adj16FinishArgs f)
{
    __shared__ double s_red[4][12];
    const FactPlanDev &P = f.plan;
    const int j = blockIdx.y;
    const int K = P.K;
    const long K2 = 2L * K;
    const int q = blockIdx.x * 256 + threadIdx.x;
    const int L = f.L;
    double ph[12];
#pragma unroll
    for (int v = 0; v < 12; v++) ph[v] = 0.0;
    if (q < K) {
        if (j >= L - 1) {
#pragma unroll
            for (int p = 0; p < 4; p++) {
                f.g_interp[((long)p * L + j) * K2 + q] = 0.0;
                f.g_interp[((long)p * L + j) * K2 + K + q] = 0.0;
            }
        } else {
            const double hw = f.h_t[j + 1] - f.h_t[j];
            const double *M = f.M + (long)j * f.tile;
            const int ncol = f.ncol;
            const int cp = P.mode_cell[2 * q], cm = P.mode_cell[2 * q + 1];
            double mpr[A16_MOM], mpi[A16_MOM], mmr[A16_MOM], mmi[A16_MOM];
            {
                const int r = cp / P.Nc, col = cp - r * P.Nc;
#pragma unroll
                for (int s = 0; s < A16_MOM; s++) {
                    const int R = A16_MOM * r + s;
                    mpr[s] = M[(long)(2 * R) * ncol + col];
                    mpi[s] = M[(long)(2 * R + 1) * ncol + col];
                }
            }
            double2 ym = make_double2(0.0, 0.0);
            if (cm >= 0) {
                const int r = cm / P.Nc, col = cm - r * P.Nc;
#pragma unroll
                for (int s = 0; s < A16_MOM; s++) {
                    const int R = A16_MOM * r + s;
                    mmr[s] = M[(long)(2 * R) * ncol + col];
                    mmi[s] = M[(long)(2 * R + 1) * ncol + col];
                }
                ym = f.ylm[K + q];
            } else {
#pragma unroll
                for (int s = 0; s < A16_MOM; s++) { mmr[s] = 0.0; mmi[s] = 0.0; }
            }
            const double2 yp = f.ylm[q];
            double ar[4], ai[4];
            double hp = 1.0;
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const long base = ((long)s * L + j) * K2;
                ar[s] = f.interp[base + q] * hp;
                ai[s] = f.interp[base + K + q] * hp;
                hp *= hw;
            }
            hp = 1.0;
            const double mq = (double)f.m[q], kq = (double)f.k[q], nq = (double)f.n[q];
#pragma unroll
            for (int p = 0; p < 4; p++) {
                // X+ = Y+ M+_p, X- = Y- M-_p
                const double xpr = yp.x * mpr[p] - yp.y * mpi[p], xpi = yp.x * mpi[p] + yp.y * mpr[p];
                const double xmr = ym.x * mmr[p] - ym.y * mmi[p], xmi = ym.x * mmi[p] + ym.y * mmr[p];
                f.g_interp[((long)p * L + j) * K2 + q] = hp * (xpr + xmr);
                f.g_interp[((long)p * L + j) * K2 + K + q] = hp * (xmi - xpi);
                // sum_s a_s M+_{p+s},  sum_s conj(a_s) M-_{p+s}
                double spr = 0.0, spi = 0.0, smr = 0.0, smi = 0.0;
#pragma unroll
                for (int s = 0; s < 4; s++) {
                    spr += ar[s] * mpr[p + s] - ai[s] * mpi[p + s];
                    spi += ar[s] * mpi[p + s] + ai[s] * mpr[p + s];
                    smr += ar[s] * mmr[p + s] + ai[s] * mmi[p + s];
                    smi += ar[s] * mmi[p + s] - ai[s] * mmr[p + s];
                }
                const double P_ = (yp.x * spi + yp.y * spr) - (ym.x * smi + ym.y * smr);
                ph[3 * p + 0] = mq * P_;
                ph[3 * p + 1] = kq * P_;
                ph[3 * p + 2] = nq * P_;
                hp *= hw;
            }
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int v = 0; v < 12; v++) {
        double x = ph[v];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) s_red[wv][v] = x;
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        const double x = (s_red[0][threadIdx.x] + s_red[1][threadIdx.x]) +
                         (s_red[2][threadIdx.x] + s_red[3][threadIdx.x]);
        f.part_phase[((long)j * gridDim.x + blockIdx.x) * 12 + threadIdx.x] = x;
    }
}

// g_phase[p][j][q] = h^p * sum over mode groups (sigma basis -> x basis).
__global__ __launch_bounds__(64) void adjoint16_phase_kernel(const double *part_phase, int groups,
                                                             const double *h_t, int L,
                                                             double *g_phase)
{
    const int j = blockIdx.x;
    const int v = threadIdx.x;
    if (v >= 12) return;
    double acc = 0.0;
    for (int b = 0; b < groups; b++) acc += part_phase[((long)j * groups + b) * 12 + v];
    const int p = v / 3, qq = v % 3;
    double hp = 1.0;
    if (j < L - 1) {
        const double hw = h_t[j + 1] - h_t[j];
        for (int k2 = 0; k2 < p; k2++) hp *= hw;
    } else {
        acc = 0.0;
    }
    g_phase[((long)p * L + j) * 3 + qq] = hp * acc;
}

int launch_adjoint16(const Adj16Args &a, int max_chunks, hipStream_t stream)
{
    if (max_chunks <= 0) return FEW_OK;
    const size_t lds = adjoint16_lds_bytes(a.RBM, a.NBC, a.plan.Nr);
    if (lds > 160 * 1024 || a.NBC < 1 || a.NBC > 8) {
        set_error("mixed adjoint: tile of %zu LDS bytes / %d column blocks unsupported", lds, a.NBC);
        return FEW_ERR_UNSUPPORTED;
    }
    switch (a.RBM) {
#define A16_CASE(X)                                                                             \
    case X:                                                                                     \
        FEW_HIP(hipFuncSetAttribute((const void *)adjoint16_kernel<X>,                          \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));     \
        hipLaunchKernelGGL(adjoint16_kernel<X>, dim3(max_chunks), dim3(64 * a.NBC), lds, stream, \
                           a);                                                                  \
        break;
        A16_CASE(1) A16_CASE(2) A16_CASE(3) A16_CASE(4) A16_CASE(5) A16_CASE(6) A16_CASE(7)
        A16_CASE(8) A16_CASE(9) A16_CASE(10) A16_CASE(11) A16_CASE(12) A16_CASE(13) A16_CASE(14)
        A16_CASE(15) A16_CASE(16) A16_CASE(17) A16_CASE(18) A16_CASE(19) A16_CASE(20) A16_CASE(21)
        A16_CASE(22) A16_CASE(23) A16_CASE(24)
#undef A16_CASE
    default:
        set_error("mixed adjoint: %d row blocks unsupported", a.RBM);
        return FEW_ERR_UNSUPPORTED;
    }
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

int launch_adjoint16_finish(const Adj16Args &a, const Adj16FinishArgs &f, hipStream_t stream)
{
    const long tile = f.tile;
    hipLaunchKernelGGL(adjoint16_reduce_kernel, dim3((unsigned)((tile + 255) / 256), a.nseg), dim3(256),
                       0, stream, a.part, a.prefix, a.nseg, tile, f.M);
    FEW_HIP(hipGetLastError());
    const int groups = (f.plan.K + 255) / 256;
    hipLaunchKernelGGL(adjoint16_finish_kernel, dim3(groups, f.L), dim3(256), 0, stream, f);
    FEW_HIP(hipGetLastError());
    hipLaunchKernelGGL(adjoint16_phase_kernel, dim3(f.L), dim3(64), 0, stream, f.part_phase, groups,
                       f.h_t, f.L, f.g_phase);
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

}  // namespace few

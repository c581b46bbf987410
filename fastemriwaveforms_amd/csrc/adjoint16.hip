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
// NBC * NH waves (16-column block cb of n, every NH-th 16-row block).  Per
// 16-sample step the workgroup builds in LDS, one job per thread, the split
// A tile (rows 2R, 2R+1 of complex row R = 7 r + s, a = sigma^s * scale *
// conj(g) * u_r; a job is one (sample, r)) and the split v table (NBC*16
// columns x 16 samples; a job is one (sample, 8 columns)); then every wave
// runs its (RBM / NH) x 3 MFMAs.  Phases are FP64 (turns, reduced), sincos is
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
#define A16_VG 8               // v-table columns per job: one sincos, then the z recurrence
#define INV_2PI_A 0.15915494309189533577

struct A16Layout {
    size_t a_hi, a_lo, v_hi, v_lo, pp, pt, pr, sig, cg, z, ph, red, total;
};

static __host__ __device__ inline size_t a16_align(size_t x) { return (x + 15) & ~(size_t)15; }

__host__ __device__ inline A16Layout a16_layout(int RBM, int NBC, int /*Nr*/)
{
    A16Layout L;
    size_t o = 0;
    L.a_hi = o; o = a16_align(o + (size_t)RBM * 16 * A16_AST * 2);
    L.a_lo = o; o = a16_align(o + (size_t)RBM * 16 * A16_AST * 2);
    L.v_hi = o; o = a16_align(o + (size_t)NBC * 16 * A16_VST * 2);
    L.v_lo = o; o = a16_align(o + (size_t)NBC * 16 * A16_VST * 2);
    L.pp = o;   o = a16_align(o + (size_t)A16_SUPER * 8);
    L.pt = o;   o = a16_align(o + (size_t)A16_SUPER * 8);
    L.pr = o;   o = a16_align(o + (size_t)A16_SUPER * 8);
    L.sig = o;  o = a16_align(o + (size_t)A16_SUPER * 4);
    L.cg = o;   o = a16_align(o + (size_t)A16_SUPER * 8);
    L.z = o;    o = a16_align(o + (size_t)A16_SUPER * 16);
    L.ph = o;   o = a16_align(o + 24 * 8);
    L.red = o;  o = a16_align(o + 16 * 4);
    L.total = o;
    return L;
}

size_t adjoint16_lds_bytes(int RBM, int NBC, int Nr) { return a16_layout(RBM, NBC, Nr).total; }

// Waves per 16-column block: two (each taking every other row block) while
// the workgroup stays within 8 waves, else one.
static inline int a16_nh(int NBC) { return NBC <= 4 ? 2 : 1; }

// hi/lo FP16 parts of a pair of floats (one packed conversion each)
__device__ __forceinline__ void split_pair(v2f x, v2h &h, v2h &l)
{
    h = __builtin_convertvector(x, v2h);
    l = f16_split_lo(h, x);
}

// One workgroup = one chunk of samples.  Per 16-sample step every thread
// takes one job -- (sample, row r): u_r and the 7 moment rows of the A tile;
// or (sample, 8 columns): the split v table -- then each wave (column block
// cb, row parity `part`) runs its MFMAs.  Two barriers per step.
template <int RBM, int NH>
__global__ __launch_bounds__(512) void adjoint16_kernel(Adj16Args a)
{
    constexpr int RBH = (RBM + NH - 1) / NH;  // row blocks per wave
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const FactPlanDev &P = a.plan;
    const int NBC = a.NBC;
    const int NW = NBC * NH;
    const int NT = 64 * NW;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6, li = lane & 15, g = lane >> 4;
    const int cb = wv % NBC, part = wv / NBC;
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
    double *sPp = reinterpret_cast<double *>(smem + Lay.pp);
    double *sPt = reinterpret_cast<double *>(smem + Lay.pt);
    double *sPr = reinterpret_cast<double *>(smem + Lay.pr);
    float *sSig = reinterpret_cast<float *>(smem + Lay.sig);
    float2 *sCg = reinterpret_cast<float2 *>(smem + Lay.cg);
    double2 *sZ = reinterpret_cast<double2 *>(smem + Lay.z);
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
    for (int w = 0; w < NW; w++) gmax = fmaxf(gmax, sRed[w]);
    int ex = 0;
    if (gmax > 0.f) frexpf(gmax, &ex);  // gmax in [2^(ex-1), 2^ex)
    const float scale = ldexpf(1.f, 13 - ex);
    const float inv_out = ldexpf(1.f, ex - 13 - 8);  // 1 / (scale * 256)

    const double inv_pw = 1.0 / sg.pw;
    const double inv_hw = 1.0 / sg.hw;
    v4f acc[RBH];
#pragma unroll
    for (int k = 0; k < RBH; k++) acc[k] = v4f{0.f, 0.f, 0.f, 0.f};

    // job slots: (sample, row) jobs, then (sample, column group) jobs, each
    // kind starting on a wave boundary so no wave mixes the two
    const int u_slots = ((Nr * 16 + 63) >> 6) << 6;
    const int v_jobs = (NBC * 16 / A16_VG) * 16;
    const int n_slots = u_slots + (((v_jobs + 63) >> 6) << 6);

    for (int sbase = lo; sbase < hi; sbase += A16_SUPER) {
        __syncthreads();
        for (int q = tid; q < A16_SUPER; q += NT) {
            const int i = sbase + q;
            const double t = a.dt * (double)i;
            const double s = (t - sg.pt) * inv_pw;
            sPp[q] = dense_phase(sPh, s, 1.0 - s) * INV_2PI_A;
            sPt[q] = P.has_k ? dense_phase(sPh + 8, s, 1.0 - s) * INV_2PI_A : 0.0;
            const double pr = dense_phase(sPh + 16, s, 1.0 - s) * INV_2PI_A;
            sPr[q] = pr;
            double zs, zc;
            sincos_turns_short(reduce_turns(pr), zs, zc);
            sZ[q] = make_double2(zc, -zs);  // exp(-i Phi_r)
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
            for (int slot = tid; slot < n_slots; slot += NT) {
                if (slot < u_slots) {
                    // A tile rows of complex row R = 7 r + s: a = sigma^s scale conj(g) u_r;
                    // each (re, im) pair of a row is one 32-bit LDS store
                    const int r = slot >> 4, ii = slot & 15;
                    if (r < Nr) {
                        const int2 mk = P.row_mk[r];
                        double sn, cs;
                        sincos_turns_short(reduce_turns((double)mk.x * sPp[i0 + ii] +
                                                  (double)mk.y * sPt[i0 + ii]),
                                     sn, cs);
                        const float ux = (float)cs, uy = (float)(-sn);
                        const float sg1 = sSig[i0 + ii];
                        const float2 cgv = sCg[i0 + ii];
                        float pw = 1.f;
#pragma unroll
                        for (int sm = 0; sm < A16_MOM; sm++) {
                            const float wx = pw * cgv.x, wy = pw * cgv.y;
                            pw *= sg1;
                            const v2f av = {wx * ux - wy * uy, wx * uy + wy * ux};
                            v2h h, l;
                            split_pair(av, h, l);
                            const int o0 = (2 * (A16_MOM * r + sm)) * A16_AST + 2 * ii;
                            const int o1 = o0 + A16_AST;
                            *reinterpret_cast<v2h *>(sAh + o0) = v2h{h.x, -h.y};
                            *reinterpret_cast<v2h *>(sAh + o1) = v2h{h.y, h.x};
                            *reinterpret_cast<v2h *>(sAl + o0) = v2h{l.x, -l.y};
                            *reinterpret_cast<v2h *>(sAl + o1) = v2h{l.y, l.x};
                        }
                    }
                } else {
                    // v_n(t_i) * 256, split, for A16_VG consecutive columns
                    const int e = slot - u_slots;
                    if (e < v_jobs) {
                        const int ii = e & 15, grp = e >> 4;
                        double sn, cs;
                        sincos_turns_short(reduce_turns((double)(P.n_lo + A16_VG * grp) * sPr[i0 + ii]), sn,
                                     cs);
                        double vr = 256.0 * cs, vi = -256.0 * sn;
                        const double2 z = sZ[i0 + ii];
#pragma unroll
                        for (int t2 = 0; t2 < A16_VG; t2++) {
                            const int col = A16_VG * grp + t2;
                            const bool in = col < P.Nc;
                            const v2f xv = {in ? f32_opaque((float)vr) : 0.f,
                                            in ? f32_opaque((float)vi) : 0.f};
                            v2h h, l;
                            split_pair(xv, h, l);
                            const int o = col * A16_VST + 2 * ii;
                            *reinterpret_cast<v2h *>(sVh + o) = h;
                            *reinterpret_cast<v2h *>(sVl + o) = l;
                            if (t2 + 1 < A16_VG) {
                                const double nr = vr * z.x - vi * z.y;
                                const double ni = vr * z.y + vi * z.x;
                                vr = nr;
                                vi = ni;
                            }
                        }
                    }
                }
            }
            __syncthreads();
            const v8h bh = *reinterpret_cast<const v8h *>(sVh + (16 * cb + li) * A16_VST + 8 * g);
            const v8h bl = *reinterpret_cast<const v8h *>(sVl + (16 * cb + li) * A16_VST + 8 * g);
            const _Float16 *aH = sAh + li * A16_AST + 8 * g;
            const _Float16 *aL = sAl + li * A16_AST + 8 * g;
#pragma unroll
            for (int k = 0; k < RBH; k++) {
                const int rb = part + NH * k;
                if (RBM % NH == 0 || rb < RBM) {
                    const v8h ah = *reinterpret_cast<const v8h *>(aH + rb * 16 * A16_AST);
                    const v8h al = *reinterpret_cast<const v8h *>(aL + rb * 16 * A16_AST);
                    acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, acc[k], 0, 0, 0);
                    acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, acc[k], 0, 0, 0);
                    acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc[k], 0, 0, 0);
                }
            }
        }
    }
    // chunk moments: D[16 rb + 4 g + v][16 cb + li] / (scale * 256)
    const int ncol = NBC * 16;
    float *out = a.part + (long)c * RBM * 16 * ncol;
#pragma unroll
    for (int k = 0; k < RBH; k++) {
        const int rb = part + NH * k;
        if (RBM % NH == 0 || rb < RBM) {
#pragma unroll
            for (int v = 0; v < 4; v++)
                out[(long)(16 * rb + 4 * g + v) * ncol + 16 * cb + li] = acc[k][v] * inv_out;
        }
    }
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
    const int nh = a16_nh(a.NBC);
    const dim3 block(64 * a.NBC * nh);
    switch (a.RBM) {
#define A16_CASE(X)                                                                             \
    case X:                                                                                     \
        if (nh == 2) {                                                                          \
            FEW_LDS((adjoint16_kernel<X, 2>), lds);                                                \
            hipLaunchKernelGGL((adjoint16_kernel<X, 2>), dim3(max_chunks), block, lds, stream, a); \
        } else {                                                                                \
            FEW_LDS((adjoint16_kernel<X, 1>), lds);                                                \
            hipLaunchKernelGGL((adjoint16_kernel<X, 1>), dim3(max_chunks), block, lds, stream, a); \
        }                                                                                       \
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

int launch_adjoint16_finish(const Adj16Args &a, const Adj16FinishArgs &f, hipStream_t stream,
                            bool reduce_chunks)
{
    const long tile = f.tile;
    if (reduce_chunks && a.nseg > 0) {
        hipLaunchKernelGGL(adjoint16_reduce_kernel, dim3((unsigned)((tile + 255) / 256), a.nseg),
                           dim3(256), 0, stream, a.part, a.prefix, a.nseg, tile, f.M);
        FEW_HIP(hipGetLastError());
    }
    const int groups = (f.plan.K + 255) / 256;
    hipLaunchKernelGGL(adjoint16_finish_kernel, dim3(groups, f.L), dim3(256), 0, stream, f);
    FEW_HIP(hipGetLastError());
    hipLaunchKernelGGL(adjoint16_phase_kernel, dim3(f.L), dim3(64), 0, stream, f.part_phase, groups,
                       f.h_t, f.L, f.g_phase);
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

}  // namespace few

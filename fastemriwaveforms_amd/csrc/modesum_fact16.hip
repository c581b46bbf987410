// Factorised mode sum on FP16 MFMA with split operands ("mixed16").
//
// Same algebra as modesum_fact.hip (W[(r,p), t] = sum_n C_p[r, n] v_n(t),
// h(t) = sum s^p u_r W), but the n-contraction runs on
// v_mfma_f32_16x16x32_f16 with each operand split into two FP16 parts,
//     A = A_hi + A_lo,  B = B_hi + B_lo   (A_hi = fp16(A), A_lo = fp16(A - A_hi)),
// and three products A_lo B_hi + A_hi B_lo + A_hi B_hi accumulated in FP32
// (the dropped A_lo B_lo is below 2^-22 relative).  The rows of each power p
// of A are scaled by one power of two so their largest entry sits near 2^14
// (no FP16 overflow; FP16 keeps its relative precision below that); B =
// 256 v_n(t).
// The scales are undone exactly in the epilogue.  Phases stay FP64 and the
// output accumulates in FP64, as in the FP32 kernel.
//
// Tile layout (per knot interval, identical in global memory and LDS):
//   plane hi, plane lo: RB*16 real rows x 32*KB real columns, real row
//       2R+e, real column k = 2a+f, holding [[Re, -Im], [Im, Re]] of
//       C[R, n_a] (against B rows Re v_n, Im v_n of tile column a).  When the
//       plan folds (n range about symmetric), tile column a is n = a >= 0 and
//       also carries D = C[R, -a]: as v_{-n} = conj(v_n), the entries are
//       [[Cr + Dr, Di - Ci], [Ci + Di, Cr - Dr]], which halves the columns;
//       stored as KB x RB blocks of 16 rows x 32 halfs (kb-major,
//       one KiB each) so that every fragment address inside the K loop is a
//       lane base plus a compile-time offset; the four 16-byte chunks of a
//       row are XOR-swizzled by (row >> 1) & 3.  A ds_read_b128 is served
//       eight lanes at a time from 32 four-byte banks (128 bytes); rows are
//       64 bytes, so the eight rows of such a lane group fall on two bank
//       halves, and the swizzle gives the four row pairs four different
//       chunks: eight distinct 16-byte bank slots, no conflict (the previous
//       (row >> 2) & 3 swizzle left a 2-way conflict on every read, measured
//       with SQ_LDS_BANK_CONFLICT).  The A fragment of a lane is 8
//       consecutive halfs (one ds_read_b128);
//   inv_sigma: 4 floats, 1/scale of the rows of power p = 0..3.

#include <math.h>

#include "batch.hpp"
#include "few_math.hpp"
#include "modesum_direct.hpp"
#include "modesum_fact.hpp"

namespace few {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

// Halfs per 16 x 32 block of a tile plane.
constexpr int F16_BLK = 512;

// Offset (in halfs) of real row rho, real column k inside one tile plane.
__device__ __forceinline__ long f16_tile_off(int rho, int k, int RB)
{
    const int rr = rho & 15;
    const int chunk = ((k & 31) >> 3) ^ ((rr >> 1) & 3);
    return ((long)((k >> 5) * RB + (rho >> 4)) * 16 + rr) * 32 + chunk * 8 + (k & 7);
}

// Lane part of a fragment read: row li of the block, 16-byte chunk g.
__device__ __forceinline__ int f16_lane_off(int li, int g) { return li * 32 + ((g ^ ((li >> 1) & 3)) << 3); }

#define INV_2PI16 0.15915494309189533577

// ------------------------------------------------------------- prep kernel
// One 512-thread workgroup per interval: each thread gathers its cells
// (fact_cell: C_p[r, n], FP64, fixed order per cell, h^p applied), the
// workgroup reduces the largest entry of each power, each power's rows are
// scaled by one power of two (exact; the largest entry near 2^14, so no FP16
// overflow) and the split FP16 planes are written.  The gathers are the
// kernel's HBM traffic: up to 4 cells per thread are kept in registers
// between the reduction and the split, larger tiles gather twice.

__device__ __forceinline__ double f16_cell_max(const double (&x)[4])
{
    return fmax(fmax(fabs(x[0]), fabs(x[1])), fmax(fabs(x[2]), fabs(x[3])));
}

// The split of one cell's powers into the two planes (rows 2R and 2R + 1 of
// R = 4 r + p, columns 2kc and 2kc + 1).  The two halfs of a row's column pair
// are adjacent (even offset): one 32-bit store each.
__device__ __forceinline__ void f16_store_cell(_Float16 *hi, _Float16 *lo, int RB, int r, int kc,
                                               const double (&x)[4][4], const double (&sigma)[4])
{
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const int R = 4 * r + p;
        const v2f a = {f32_opaque((float)(x[p][0] * sigma[p])), f32_opaque((float)(x[p][1] * sigma[p]))};
        const v2f b = {f32_opaque((float)(x[p][2] * sigma[p])), f32_opaque((float)(x[p][3] * sigma[p]))};
        const v2h ah = __builtin_convertvector(a, v2h), bh = __builtin_convertvector(b, v2h);
        const long o0 = f16_tile_off(2 * R, 2 * kc, RB);
        const long o1 = f16_tile_off(2 * R + 1, 2 * kc, RB);
        *reinterpret_cast<v2h *>(hi + o0) = ah;
        *reinterpret_cast<v2h *>(lo + o0) = f16_split_lo(ah, a);
        *reinterpret_cast<v2h *>(hi + o1) = bh;
        *reinterpret_cast<v2h *>(lo + o1) = f16_split_lo(bh, b);
    }
}

#define F16_PREP_THREADS 512

// The interval's scales from the per-wave maxima in s_red; writes the inverse
// scales after the planes (thread 0).
__device__ __forceinline__ void f16_interval_sigma(const double (*s_red)[4], int nwaves,
                                                   float *inv_out, double (&sigma)[4])
{
#pragma unroll
    for (int p = 0; p < 4; p++) {
        double m = 0.0;
        for (int w = 0; w < nwaves; w++) m = fmax(m, s_red[w][p]);
        int ex = 0;
        if (m > 0.0) {
            frexp(m, &ex);
            ex = 14 - ex;
        }
        sigma[p] = ldexp(1.0, ex);
        if (threadIdx.x == 0) inv_out[p] = (float)ldexp(1.0, -ex);
    }
}

// One workgroup per interval with the cells gathered once: each thread keeps
// the CPT cells it gathered (4 powers, FP64) in registers through the maxima
// reduction and splits them from there (the two-pass form gathers every cell
// twice; the gathers are the kernel's HBM traffic).
template <int CPT>
__device__ void fact16_prep_interval_cached(const FactPlanDev &P, const double *interp,
                                            const double2 *ylm, const double *h_t, int L, int j,
                                            unsigned char *tile, int KB, int RB)
{
    __shared__ double s_red[F16_PREP_THREADS / 64][4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int ncolc = 16 * KB;
    const int ncell = 2 * RB * ncolc;
    const double hw = h_t[j + 1] - h_t[j];
    double x[CPT][4][4];
    double mx[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < CPT; k++) {
        const int ci = tid + k * F16_PREP_THREADS;
        const int r = ci / ncolc, col = ci - r * ncolc;
        if (ci < ncell) {
            fact_cell(P, interp, ylm, L, j, hw, r, col, x[k]);
        } else {
#pragma unroll
            for (int p = 0; p < 4; p++) x[k][p][0] = x[k][p][1] = x[k][p][2] = x[k][p][3] = 0.0;
        }
#pragma unroll
        for (int p = 0; p < 4; p++) mx[p] = fmax(mx[p], f16_cell_max(x[k][p]));
    }
#pragma unroll
    for (int p = 0; p < 4; p++) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mx[p] = fmax(mx[p], __shfl_xor(mx[p], off, 64));
        if (lane == 0) s_red[wv][p] = mx[p];
    }
    __syncthreads();
    const long plane = (long)KB * RB * F16_BLK;
    _Float16 *hi = reinterpret_cast<_Float16 *>(tile);
    _Float16 *lo = hi + plane;
    double sigma[4];
    f16_interval_sigma(s_red, F16_PREP_THREADS / 64, reinterpret_cast<float *>(lo + plane), sigma);
#pragma unroll
    for (int k = 0; k < CPT; k++) {
        const int ci = tid + k * F16_PREP_THREADS;
        if (ci < ncell) f16_store_cell(hi, lo, RB, ci / ncolc, ci % ncolc, x[k], sigma);
    }
}

// Cells per thread of an interval's tile (the kernels are specialised on it, so
// the common small tiles get the registers of their own path: two workgroups
// per CU for up to two cells per thread); 0: more than 4, gathered twice.
inline int f16_prep_cpt(int KB, int RB)
{
    const int cpt = (2 * RB * 16 * KB + F16_PREP_THREADS - 1) / F16_PREP_THREADS;
    return cpt <= 4 ? (cpt < 1 ? 1 : cpt) : 0;
}

template <int CPT>
__device__ void fact16_prep_interval(const FactPlanDev &P, const double *interp, const double2 *ylm,
                                     const double *h_t, int L, int j, unsigned char *tile, int KB,
                                     int RB)
{
    if constexpr (CPT > 0) {
        return fact16_prep_interval_cached<CPT>(P, interp, ylm, h_t, L, j, tile, KB, RB);
    }
    __shared__ double s_red[F16_PREP_THREADS / 64][4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int ncolc = 16 * KB;         // complex columns incl. padding
    const int ncell = 2 * RB * ncolc;  // complex rows r incl. padding x columns
    const double hw = h_t[j + 1] - h_t[j];
    double mx[4] = {0.0, 0.0, 0.0, 0.0};
    for (int ci = tid; ci < ncell; ci += F16_PREP_THREADS) {
        const int r = ci / ncolc, col = ci - r * ncolc;
        double x[4][4];
        fact_cell(P, interp, ylm, L, j, hw, r, col, x);
#pragma unroll
        for (int p = 0; p < 4; p++) mx[p] = fmax(mx[p], f16_cell_max(x[p]));
    }
#pragma unroll
    for (int p = 0; p < 4; p++) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mx[p] = fmax(mx[p], __shfl_xor(mx[p], off, 64));
        if (lane == 0) s_red[wv][p] = mx[p];
    }
    __syncthreads();
    const long plane = (long)KB * RB * F16_BLK;
    _Float16 *hi = reinterpret_cast<_Float16 *>(tile);
    _Float16 *lo = hi + plane;
    double sigma[4];
    f16_interval_sigma(s_red, F16_PREP_THREADS / 64, reinterpret_cast<float *>(lo + plane), sigma);
    for (int ci = tid; ci < ncell; ci += F16_PREP_THREADS) {
        const int r = ci / ncolc, col = ci - r * ncolc;
        double x[4][4];
        fact_cell(P, interp, ylm, L, j, hw, r, col, x);
        f16_store_cell(hi, lo, RB, r, col, x, sigma);
    }
}

// Two resident workgroups per CU where at most two cells per thread are held.
#define F16_PREP_BOUNDS(CPT) __launch_bounds__(F16_PREP_THREADS, (CPT) >= 1 && (CPT) <= 2 ? 2 : 1)

template <int CPT>
__global__ F16_PREP_BOUNDS(CPT) void fact16_prep_whole_kernel(Fact16PrepArgs a)
{
    const int j = blockIdx.x;
    fact16_prep_interval<CPT>(a.plan, a.interp, a.ylm, a.h_t, a.L, j,
                              a.tiles + (long)j * a.tile_bytes, a.KB, a.RB);
}

int launch_fact16_prep(const Fact16PrepArgs &a, hipStream_t stream)
{
    if (a.L < 2) return FEW_OK;
    const dim3 grid(a.L - 1), block(F16_PREP_THREADS);
    switch (f16_prep_cpt(a.KB, a.RB)) {
    case 1: hipLaunchKernelGGL(fact16_prep_whole_kernel<1>, grid, block, 0, stream, a); break;
    case 2: hipLaunchKernelGGL(fact16_prep_whole_kernel<2>, grid, block, 0, stream, a); break;
    case 3: hipLaunchKernelGGL(fact16_prep_whole_kernel<3>, grid, block, 0, stream, a); break;
    case 4: hipLaunchKernelGGL(fact16_prep_whole_kernel<4>, grid, block, 0, stream, a); break;
    default: hipLaunchKernelGGL(fact16_prep_whole_kernel<0>, grid, block, 0, stream, a); break;
    }
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

// ------------------------------------------------------------- sum kernel
// v_n recurrence state type of the K loop: FP32 (the operands are split to
// FP16 hi/lo pairs, ~2^-22 relative, right after; three FP32 steps by z^16,
// itself rounded from FP64, add a few 2^-24)
typedef float f16v_t;

__device__ __forceinline__ void split_pack(const f16v_t (&vr)[4], const f16v_t (&vi)[4], v8h &bh,
                                           v8h &bl)
{
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const v2f x = {(float)vr[t], (float)vi[t]};
        const v2h h = __builtin_convertvector(x, v2h);  // one packed conversion
        const v2h l = f16_split_lo(h, x);
        bh[2 * t] = h.x;
        bh[2 * t + 1] = h.y;
        bl[2 * t] = l.x;
        bl[2 * t + 1] = l.y;
    }
}

template <typename T>
__device__ __forceinline__ void cmul(T ar, T ai, T br, T bi, T &cr, T &ci)
{
    cr = ar * br - ai * bi;
    ci = ar * bi + ai * br;
}

constexpr int F16_NB = 2;        // 16-sample column blocks per wave

// WAVES waves per workgroup share one LDS tile; 4 waves run one wave per SIMD
// with the full register file, 8 waves two per SIMD.
constexpr int f16_tile_samples(int waves) { return 16 * F16_NB * waves; }

// (ar + i ai)(br + i bi) in FP32 with the fused operations spelled out, so
// every kernel that inlines it rounds the same way.
__device__ __forceinline__ void cmul_f32(float ar, float ai, float br, float bi, float &cr, float &ci)
{
    cr = fmaf(ar, br, -(ai * bi));
    ci = fmaf(ar, bi, ai * br);
}

// v_n -> v_n z^16 for the next 32-column step (FP32 complex products).
__device__ __forceinline__ void f16_advance(f16v_t (&vr)[F16_NB][4], f16v_t (&vi)[F16_NB][4],
                                            const f16v_t (&zr16)[F16_NB],
                                            const f16v_t (&zi16)[F16_NB])
{
#pragma unroll
    for (int c = 0; c < F16_NB; c++)
#pragma unroll
        for (int t2 = 0; t2 < 4; t2++) {
            f16v_t nr, ni;
            cmul_f32(vr[c][t2], vi[c][t2], zr16[c], zi16[c], nr, ni);
            vr[c][t2] = nr;
            vi[c][t2] = ni;
        }
}

// One 32-column step of the n-contraction: B fragments from v_n (then v_n
// advanced by z^16 when another step follows), A fragments from the tile
// block column at aH / aL, and the three split products per row block.
template <int RB, bool FIRST>
__device__ __forceinline__ void f16_kstep(const _Float16 *aH, const _Float16 *aL, bool advance,
                                          f16v_t (&vr)[F16_NB][4], f16v_t (&vi)[F16_NB][4],
                                          const f16v_t (&zr16)[F16_NB],
                                          const f16v_t (&zi16)[F16_NB], v4f (&acc)[RB][F16_NB])
{
    v8h bh[F16_NB], bl[F16_NB];
#pragma unroll
    for (int c = 0; c < F16_NB; c++) split_pack(vr[c], vi[c], bh[c], bl[c]);
    if (advance) f16_advance(vr, vi, zr16, zi16);
    // A fragments of the next row block are loaded while the current one
    // feeds its MFMAs; the scheduling barrier keeps the loads from being
    // hoisted further (register pressure).
    v8h ahn = *reinterpret_cast<const v8h *>(aH);
    v8h aln = *reinterpret_cast<const v8h *>(aL);
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
        const v8h ah = ahn, al = aln;
        if (rb + 1 < RB) {
            ahn = *reinterpret_cast<const v8h *>(aH + (rb + 1) * F16_BLK);
            aln = *reinterpret_cast<const v8h *>(aL + (rb + 1) * F16_BLK);
        }
#pragma unroll
        for (int c = 0; c < F16_NB; c++)
            acc[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                al, bh[c], FIRST ? v4f{0.f, 0.f, 0.f, 0.f} : acc[rb][c], 0, 0, 0);
#pragma unroll
        for (int c = 0; c < F16_NB; c++)
            acc[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl[c], acc[rb][c], 0, 0, 0);
#pragma unroll
        for (int c = 0; c < F16_NB; c++)
            acc[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh[c], acc[rb][c], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
    }
}

// One tile (tile q of the interval whose coefficient tile and phase
// coefficients sit in LDS) of the FP16 split-operand factorised sum runs in
// three phases: setup (FP64 phases, v_n recurrence state, prefetch of the
// output samples), contraction (MFMA), epilogue (row/power sum in FP64 and
// the output update).

// v_n recurrence state of one wave between setup and contraction
struct F16V {
    f16v_t vr[F16_NB][4], vi[F16_NB][4], zr16[F16_NB], zi16[F16_NB];
    // arithmetic rows: the sample phasor this lane group evaluated at setup
    // (sin, cos), handed to the other lane groups of the sample in the epilogue
    double ps[F16_NB], pc[F16_NB];
};

// First output sample of a lane in tile q of an interval.
template <int WAVES>
__device__ __forceinline__ int f16_base(const Segment &sg, int q)
{
    const int tid = threadIdx.x;
    return sg.start + q * f16_tile_samples(WAVES) + (tid >> 6) * (16 * F16_NB) + (tid & 15);
}

// The output samples a tile adds into, fetched at setup so the HBM latency
// hides behind the contraction (lane group 0 owns the update).
__device__ __forceinline__ void f16_prefetch(const Segment &sg, int base_i, const double2 *wave,
                                             double2 (&wold)[F16_NB], bool overwrite = false)
{
    const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int c = 0; c < F16_NB; c++) {
        const int idx = base_i + c * 16;
        wold[c] = (!overwrite && g == 0 && idx < sg.end) ? wave[idx] : make_double2(0.0, 0.0);
    }
}

// Phi_r (FP64) and the v_n recurrence state: lane group g needs v_n for
// n = n0 + 16 kb + 4 g + t, t = 0..3, scaled by 256 (n0: the n of tile column 0).
template <bool ARITH>
__device__ __forceinline__ void f16_setup(const FactPlanDev &P, const double *sPh, const Segment &sg,
                                          double inv_pw, int base_i, double dt, F16V &V)
{
    const int lane = threadIdx.x & 63, g = lane >> 4, li = lane & 15;
#pragma unroll
    for (int c = 0; c < F16_NB; c++) {
        const double t = dt * (double)(base_i + c * 16);
        const double s = (t - sg.pt) * inv_pw;
        double ur, zsn, zcs;
        if constexpr (ARITH) {
            // The four lane groups of a sample share its phases and phasors:
            // group g evaluates one of them and the others read it (ds_bpermute)
            //   g = 0:     Phi_r             -> z = exp(-i Phi_r)
            //   g = 1:     Phi_phi           -> e = exp(-2 i Phi_phi)      (epilogue)
            //   g = 2, 3:  Phi_phi, Phi_theta -> u = exp(-i((m0 + g - 2) Phi_phi
            //                                               + k0 Phi_theta)) (epilogue)
            const double p1 = dense_phase_fused(sPh + (g == 0 ? 16 : 0), s, 1.0 - s) * INV_2PI16;
            const double p2 = P.has_k ? dense_phase_fused(sPh + 8, s, 1.0 - s) * INV_2PI16 : 0.0;
            const double arg = g == 0 ? p1
                             : g == 1 ? 2.0 * p1
                                      : (double)(P.m0 + g - 2) * p1 + (double)P.k0 * p2;
            double sn, cs;
            sincos_turns_circle(reduce_turns(arg), sn, cs);
            V.ps[c] = sn;
            V.pc[c] = cs;
            ur = __shfl(p1, li);
            zsn = __shfl(sn, li);
            zcs = __shfl(cs, li);
        } else {
            ur = dense_phase_fused(sPh + 16, s, 1.0 - s) * INV_2PI16;
            sincos_turns_circle(reduce_turns(ur), zsn, zcs);
        }
        // lane group g needs v_n for n = n0 + 16 kb + 4 g + t, t = 0..3 (scaled by 256)
        const double zr = zcs, zi = -zsn;  // z = exp(-i Phi_r)
        double z2r, z2i, z4r, z4i, z8r, z8i, z16r, z16i;  // powers of z by squaring
        cmul(zr, zi, zr, zi, z2r, z2i);
        cmul(z2r, z2i, z2r, z2i, z4r, z4i);
        cmul(z4r, z4i, z4r, z4i, z8r, z8i);
        cmul(z8r, z8i, z8r, z8i, z16r, z16i);
        double vr, vi;
        if (P.tn0 == 0) {
            // v_{4g} = z^{4g} = (z^4 if g odd) (z^8 if g >= 2): no second sincos
            const double ar = (g & 1) ? 256.0 * z4r : 256.0, ai = (g & 1) ? 256.0 * z4i : 0.0;
            const double br = (g & 2) ? z8r : 1.0, bi = (g & 2) ? z8i : 0.0;
            cmul(ar, ai, br, bi, vr, vi);
        } else {
            double sn, cs;
            sincos_turns_circle(reduce_turns((double)(P.tn0 + 4 * g) * ur), sn, cs);
            vr = 256.0 * cs;
            vi = -256.0 * sn;
        }
        V.vr[c][0] = (f16v_t)vr;
        V.vi[c][0] = (f16v_t)vi;
#pragma unroll
        for (int t2 = 1; t2 < 4; t2++) {
            double nr, ni;
            cmul(vr, vi, zr, zi, nr, ni);
            vr = nr;
            vi = ni;
            V.vr[c][t2] = (f16v_t)vr;
            V.vi[c][t2] = (f16v_t)vi;
        }
        V.zr16[c] = (f16v_t)z16r;
        V.zi16[c] = (f16v_t)z16i;
    }
}

// n-contraction of one tile on MFMA: acc = W[(r, p), samples].
template <int RB>
__device__ __forceinline__ void f16_contract(const unsigned char *sTile, int KB, F16V &V,
                                             v4f (&acc)[RB][F16_NB])
{
    const int lane = threadIdx.x & 63;
    const long plane = (long)KB * RB * F16_BLK;
    const _Float16 *sHi = reinterpret_cast<const _Float16 *>(sTile);
    const _Float16 *sLo = sHi + plane;
    const int lo_lane = f16_lane_off(lane & 15, lane >> 4);
    // kb = 0 starts the accumulators from zero (no separate clearing)
    f16_kstep<RB, true>(sHi + lo_lane, sLo + lo_lane, KB > 1, V.vr, V.vi, V.zr16, V.zi16, acc);
    // two steps per trip: the v_n state alternates between two register sets
    // instead of being copied back at every trip
    int kb = 1;
    for (; kb + 1 < KB; kb += 2) {
        f16_kstep<RB, false>(sHi + kb * RB * F16_BLK + lo_lane, sLo + kb * RB * F16_BLK + lo_lane,
                             true, V.vr, V.vi, V.zr16, V.zi16, acc);
        f16_kstep<RB, false>(sHi + (kb + 1) * RB * F16_BLK + lo_lane,
                             sLo + (kb + 1) * RB * F16_BLK + lo_lane, kb + 2 < KB, V.vr, V.vi,
                             V.zr16, V.zi16, acc);
    }
    if (kb < KB)
        f16_kstep<RB, false>(sHi + kb * RB * F16_BLK + lo_lane, sLo + kb * RB * F16_BLK + lo_lane,
                             false, V.vr, V.vi, V.zr16, V.zi16, acc);
}

// Epilogue: h = sum_{r,p} s^p u_r W / (256 sigma_R), added into the output.
template <int RB, bool ARITH>
__device__ __forceinline__ void f16_epilogue(const FactPlanDev &P, const unsigned char *sTile,
                                             int KB, const double *sPh, const Segment &sg,
                                             double inv_pw, double inv_hw, int base_i, double dt,
                                             const v4f (&acc)[RB][F16_NB], const F16V &V,
                                             double2 (&wold)[F16_NB], double2 *wave)
{
    const int g = (threadIdx.x & 63) >> 4;
    const int li = threadIdx.x & 15;
    const long plane = (long)KB * RB * F16_BLK;
    const float *sInv = reinterpret_cast<const float *>(sTile + 4 * plane);  // 1/scale per power
    const int f = g & 1;
    const int p0 = 2 * f;
#pragma unroll
    for (int c = 0; c < F16_NB; c++) {
        // sample-level quantities recomputed here rather than held across the K loop
        const int idx = base_i + c * 16;
        const double t = dt * (double)idx;
        const float s1f = (float)((t - sg.h_t) * inv_hw);
        const float s2f = s1f * s1f;
        const float spa = (p0 == 0 ? 1.f : s2f) * (1.0f / 256.0f);
        const float spb = (p0 == 0 ? s1f : s2f * s1f) * (1.0f / 256.0f);
        // powers p0, p0 + 1 of this lane's rows, with their tile scales undone
        const float ka = spa * sInv[p0];
        const float kb2 = spb * sInv[p0 + 1];
        double hre = 0.0, him = 0.0;
        if constexpr (ARITH) {
            // Rows of one lane are r = 2 rb + (g >> 1) with m_r = m_first + 2 rb:
            // sum_rb u_r z_rb = u_first * sum_rb e^rb z_rb (e = exp(-2 i Phi_phi)),
            // summed by Horner's rule in FP64.
            // e from lane group 1, u_first from group 2 + (g >> 1) (set up at setup)
            const double e_re = __shfl(V.pc[c], li + 16), e_im = -__shfl(V.ps[c], li + 16);
            // the last row block starts the sum (no multiply of a zero sum)
            hre = (double)fmaf(ka, acc[RB - 1][c][0], kb2 * acc[RB - 1][c][2]);
            him = (double)fmaf(ka, acc[RB - 1][c][1], kb2 * acc[RB - 1][c][3]);
#pragma unroll
            for (int rb = RB - 2; rb >= 0; rb--) {
                // rows r >= Nr hold zeros (W = 0)
                const v4f W = acc[rb][c];
                const double zr = (double)fmaf(ka, W[0], kb2 * W[2]);
                const double zi = (double)fmaf(ka, W[1], kb2 * W[3]);
                const double nr = fma(hre, e_re, fma(-him, e_im, zr));
                const double ni = fma(hre, e_im, fma(him, e_re, zi));
                hre = nr;
                him = ni;
            }
            const int ul = li + 32 + 16 * (g >> 1);
            const double u_re = __shfl(V.pc[c], ul), u_im = -__shfl(V.ps[c], ul);
            double nr, ni;
            cmul(u_re, u_im, hre, him, nr, ni);
            hre = nr;
            him = ni;
        } else {
            const double s = (t - sg.pt) * inv_pw;
            const double uph = dense_phase_fused(sPh, s, 1.0 - s) * INV_2PI16;
            const double uth = P.has_k ? dense_phase_fused(sPh + 8, s, 1.0 - s) * INV_2PI16 : 0.0;
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                const int r = 2 * rb + (g >> 1);
                const int2 mk = r < P.Nr ? P.row_mk[r] : make_int2(0, 0);
                double sn, cs;
                sincos_turns_circle(reduce_turns((double)mk.x * uph + (double)mk.y * uth), sn, cs);
                const float ur = (float)cs, ui = (float)(-sn);
                // rows r >= Nr hold zeros (W = 0)
                const v4f W = acc[rb][c];
                const float zr = ka * W[0] + kb2 * W[2];
                const float zi = ka * W[1] + kb2 * W[3];
                hre += (double)(ur * zr - ui * zi);
                him += (double)(ur * zi + ui * zr);
            }
        }
        hre += __shfl_xor(hre, 16);
        him += __shfl_xor(him, 16);
        hre += __shfl_xor(hre, 32);
        him += __shfl_xor(him, 32);
        wold[c].x += hre;
        wold[c].y += him;
    }
    // stores after both columns: a store between the two would make the
    // second column's wait on the prefetched samples also wait for it
#pragma unroll
    for (int c = 0; c < F16_NB; c++) {
        const int idx = base_i + c * 16;
        if (g == 0 && idx < sg.end) wave[idx] = wold[c];
    }
}

// The contraction runs at raised wave priority: of the two waves sharing a
// SIMD, the one contracting then issues its MFMAs first and the other one's
// VALU work (setup, epilogue) fills the cycles between them.
template <int RB>
__device__ __forceinline__ void f16_contract_prio(const unsigned char *sTile, int KB, F16V &V,
                                                  v4f (&acc)[RB][F16_NB], int prio)
{
    if (prio) __builtin_amdgcn_s_setprio(1);
    f16_contract<RB>(sTile, KB, V, acc);
    if (prio) __builtin_amdgcn_s_setprio(0);
}

// ------------------------------------------------ row-streamed variant
// Arithmetic rows only.  The B fragments of all KB column steps are split
// once per tile and held in registers (16 KB VGPRs); the row blocks are then
// contracted one at a time, from the last to the first, and each finished
// block is folded into the epilogue's Horner sum while the MFMAs of the next
// block run (software pipelining by one row block).  The FP64 row sum thus
// overlaps the contraction inside the wave, and only two row blocks of
// accumulators are live.  Every product, accumulation and rounding is the
// one of fact16_kernel (same MFMA chains per (row block, column block), same
// fold order), so the two give identical results.
template <int KB>
struct F16Bfrag {
    v8h bh[KB][F16_NB], bl[KB][F16_NB];
};

template <int KB>
__device__ __forceinline__ void f16_split_all(F16V &V, F16Bfrag<KB> &Bf)
{
#pragma unroll
    for (int kb = 0; kb < KB; kb++) {
#pragma unroll
        for (int c = 0; c < F16_NB; c++) split_pack(V.vr[c], V.vi[c], Bf.bh[kb][c], Bf.bl[kb][c]);
        if (kb + 1 < KB) f16_advance(V.vr, V.vi, V.zr16, V.zi16);
    }
}

// One row block: acc = W[rows of block, samples] over all KB column steps
// (aH / aL: the block's A fragment address in column step 0; kstride: the
// distance between column steps).
template <int KB>
__device__ __forceinline__ void f16_rows16(const _Float16 *aH, const _Float16 *aL, int kstride,
                                           const F16Bfrag<KB> &Bf, v4f (&acc)[F16_NB])
{
#pragma unroll
    for (int kb = 0; kb < KB; kb++) {
        const v8h ah = *reinterpret_cast<const v8h *>(aH + kb * kstride);
        const v8h al = *reinterpret_cast<const v8h *>(aL + kb * kstride);
#pragma unroll
        for (int c = 0; c < F16_NB; c++)
            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                al, Bf.bh[kb][c], kb == 0 ? v4f{0.f, 0.f, 0.f, 0.f} : acc[c], 0, 0, 0);
#pragma unroll
        for (int c = 0; c < F16_NB; c++)
            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, Bf.bl[kb][c], acc[c], 0, 0, 0);
#pragma unroll
        for (int c = 0; c < F16_NB; c++)
            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, Bf.bh[kb][c], acc[c], 0, 0, 0);
    }
}

// One Horner step of the row sum with a finished row block (the
// arithmetic-rows branch of f16_epilogue, per row block).
struct F16Fold {
    float ka[F16_NB], kb2[F16_NB];
    double er[F16_NB], ei[F16_NB], hr[F16_NB], hi[F16_NB];
};

// FIRST: the block starts the sum (as in f16_epilogue, no step from zero).
template <bool FIRST>
__device__ __forceinline__ void f16_fold(const v4f (&acc)[F16_NB], F16Fold &F)
{
#pragma unroll
    for (int c = 0; c < F16_NB; c++) {
        const v4f W = acc[c];
        const double zr = (double)fmaf(F.ka[c], W[0], F.kb2[c] * W[2]);
        const double zi = (double)fmaf(F.ka[c], W[1], F.kb2[c] * W[3]);
        if constexpr (FIRST) {
            F.hr[c] = zr;
            F.hi[c] = zi;
        } else {
            const double nr = fma(F.hr[c], F.er[c], fma(-F.hi[c], F.ei[c], zr));
            const double ni = fma(F.hr[c], F.ei[c], fma(F.hi[c], F.er[c], zi));
            F.hr[c] = nr;
            F.hi[c] = ni;
        }
    }
}

// Scheduling of one (row block, fold of the previous block) pair: per column
// step the two A-fragment reads of the next step, then the six MFMAs in pairs
// with two VALU instructions (the fold's) after each pair, so the fold issues
// in the MFMA shadow instead of after the block and the LDS reads one step
// ahead.
template <int KB>
__device__ __forceinline__ void f16s_interleave()
{
    __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);  // DS reads of step 0
#pragma unroll
    for (int kb = 0; kb < KB; kb++) {
        if (kb + 1 < KB) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);  // of step kb + 1
#pragma unroll
        for (int i = 0; i < 3; i++) {
            __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);  // MFMA
            __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);  // VALU
        }
    }
}

// One tile of the row-streamed variant: setup, B split, streamed rows with
// the folded epilogue, output update.
template <int KB>
__device__ __forceinline__ void f16s_tile(const FactPlanDev &P, const unsigned char *sTile, int RB,
                                          const double *sPh, const Segment &sg, double inv_pw,
                                          double inv_hw, int base_i, double dt, double2 *wave,
                                          bool overwrite, int prio)
{
    const int g = (threadIdx.x & 63) >> 4;
    const int li = threadIdx.x & 15;
    double2 wold[F16_NB];
    F16V V;
    f16_prefetch(sg, base_i, wave, wold, overwrite);
    f16_setup<true>(P, sPh, sg, inv_pw, base_i, dt, V);
    const int plane = KB * RB * F16_BLK;
    const float *sInv = reinterpret_cast<const float *>(sTile + 4 * (long)plane);
    const int p0 = 2 * (g & 1);
    F16Fold F;
#pragma unroll
    for (int c = 0; c < F16_NB; c++) {
        const double t = dt * (double)(base_i + c * 16);
        const float s1f = (float)((t - sg.h_t) * inv_hw);
        const float s2f = s1f * s1f;
        const float spa = (p0 == 0 ? 1.f : s2f) * (1.0f / 256.0f);
        const float spb = (p0 == 0 ? s1f : s2f * s1f) * (1.0f / 256.0f);
        F.ka[c] = spa * sInv[p0];
        F.kb2[c] = spb * sInv[p0 + 1];
        F.er[c] = __shfl(V.pc[c], li + 16);
        F.ei[c] = -__shfl(V.ps[c], li + 16);
    }
    F16Bfrag<KB> Bf;
    f16_split_all<KB>(V, Bf);
    const _Float16 *sHi = reinterpret_cast<const _Float16 *>(sTile);
    const _Float16 *aH = sHi + f16_lane_off(li, g);
    const _Float16 *aL = aH + plane;
    const int kstride = RB * F16_BLK;
    v4f accA[F16_NB], accB[F16_NB];
    if (prio) __builtin_amdgcn_s_setprio(1);
    // row blocks from the last to the first; the fold of block rb runs
    // under the MFMAs of block rb - 1
    int rb = RB - 1;
    f16_rows16<KB>(aH + rb * F16_BLK, aL + rb * F16_BLK, kstride, Bf, accA);
    if (rb == 0) {
        f16_fold<true>(accA, F);
    } else {
        f16_rows16<KB>(aH + (rb - 1) * F16_BLK, aL + (rb - 1) * F16_BLK, kstride, Bf, accB);
        f16_fold<true>(accA, F);
        f16s_interleave<KB>();
        rb -= 1;  // accB holds block rb
        while (rb >= 2) {
            f16_rows16<KB>(aH + (rb - 1) * F16_BLK, aL + (rb - 1) * F16_BLK, kstride, Bf, accA);
            f16_fold<false>(accB, F);
            f16s_interleave<KB>();
            f16_rows16<KB>(aH + (rb - 2) * F16_BLK, aL + (rb - 2) * F16_BLK, kstride, Bf, accB);
            f16_fold<false>(accA, F);
            f16s_interleave<KB>();
            rb -= 2;
        }
        if (rb == 1) {
            f16_rows16<KB>(aH, aL, kstride, Bf, accA);
            f16_fold<false>(accB, F);
            f16s_interleave<KB>();
            f16_fold<false>(accA, F);
        } else {
            f16_fold<false>(accB, F);
        }
    }
    if (prio) __builtin_amdgcn_s_setprio(0);
#pragma unroll
    for (int c = 0; c < F16_NB; c++) {
        const int ul = li + 32 + 16 * (g >> 1);
        const double u_re = __shfl(V.pc[c], ul), u_im = -__shfl(V.ps[c], ul);
        double hre, him;
        cmul(u_re, u_im, F.hr[c], F.hi[c], hre, him);
        hre += __shfl_xor(hre, 16);
        him += __shfl_xor(him, 16);
        hre += __shfl_xor(hre, 32);
        him += __shfl_xor(him, 32);
        wold[c].x += hre;
        wold[c].y += him;
    }
#pragma unroll
    for (int c = 0; c < F16_NB; c++) {
        const int idx = base_i + c * 16;
        if (g == 0 && idx < sg.end) wave[idx] = wold[c];
    }
}

// Single waveform, row-streamed variant (fact16_kernel's work split and
// tile loads; RB at run time, KB at compile time).
template <int KB, int WAVES>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(WAVES / 4, WAVES / 4)))
void fact16s_kernel(Fact16Args a, int RB)
{
    constexpr int NT = 64 * WAVES;
    extern __shared__ __attribute__((aligned(16))) unsigned char sTile[];
    const int tid = threadIdx.x;
    double *sPh = reinterpret_cast<double *>(sTile + a.tile_bytes);
    const int total = a.prefix[a.nseg];
    const int per = (total + (int)gridDim.x - 1) / (int)gridDim.x;
    const int t0 = (int)blockIdx.x * per;
    const int t1 = min(total, t0 + per);
    if (a.overwrite) {
        const long end = a.nseg > 0 ? a.segs[a.nseg - 1].end : 0;
        for (long i = end + (long)blockIdx.x * NT + tid; i < a.out_len; i += (long)gridDim.x * NT)
            a.wave[i] = make_double2(0.0, 0.0);
    }
    if (t0 >= t1) return;
    int j = upper_index(a.prefix, a.nseg, t0);
    int jnext_start = a.prefix[j + 1];
    const long np = (long)a.nseg * 3;
    bool stale = true;
    Segment sgj;
    double inv_pw = 1.0, inv_hw = 1.0;
    for (int tile = t0; tile < t1; tile++) {
        while (tile >= jnext_start) {
            j++;
            jnext_start = a.prefix[j + 1];
            stale = true;
        }
        if (stale) {
            sgj = a.segs[j];
            inv_pw = 1.0 / sgj.pw;
            inv_hw = 1.0 / sgj.hw;
            __syncthreads();
            const uint4 *src = reinterpret_cast<const uint4 *>(a.tiles + (long)j * a.tile_bytes);
            uint4 *dst = reinterpret_cast<uint4 *>(sTile);
            const int nv = (int)(a.tile_bytes >> 4);
            for (int i = tid; i < nv; i += NT) dst[i] = src[i];
            if (tid < 24) sPh[tid] = a.pcoef[(tid & 7) * np + (long)j * 3 + (tid >> 3)];
            __syncthreads();
            stale = false;
        }
        f16s_tile<KB>(a.plan, sTile, RB, sPh, sgj, inv_pw, inv_hw,
                      f16_base<WAVES>(sgj, tile - a.prefix[j]), a.dt, a.wave, a.overwrite != 0,
                      a.sched.prio);
    }
}

// Single waveform, persistent: block b owns a contiguous range of the
// (interval, tile) sequence; the LDS tile is reloaded on interval change.
template <int RB, int WAVES, bool ARITH>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(WAVES / 4, WAVES / 4)))
void fact16_kernel(Fact16Args a)
{
    constexpr int NT = 64 * WAVES;
    extern __shared__ __attribute__((aligned(16))) unsigned char sTile[];
    const int tid = threadIdx.x;
    double *sPh = reinterpret_cast<double *>(sTile + a.tile_bytes);  // 24 phase coefficients
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
    int j = upper_index(a.prefix, a.nseg, t0);
    int jnext_start = a.prefix[j + 1];
    const long np = (long)a.nseg * 3;
    bool stale = true;
    Segment sgj;
    double inv_pw = 1.0, inv_hw = 1.0;  // per interval, computed when the interval changes
    for (int tile = t0; tile < t1; tile++) {
        while (tile >= jnext_start) {
            j++;
            jnext_start = a.prefix[j + 1];
            stale = true;
        }
        if (stale) {
            sgj = a.segs[j];
            inv_pw = 1.0 / sgj.pw;
            inv_hw = 1.0 / sgj.hw;
            __syncthreads();
            const uint4 *src = reinterpret_cast<const uint4 *>(a.tiles + (long)j * a.tile_bytes);
            uint4 *dst = reinterpret_cast<uint4 *>(sTile);
            const int nv = (int)(a.tile_bytes >> 4);
            for (int i = tid; i < nv; i += NT) dst[i] = src[i];
            if (tid < 24) sPh[tid] = a.pcoef[(tid & 7) * np + (long)j * 3 + (tid >> 3)];
            __syncthreads();
            stale = false;
        }
        const int base_i = f16_base<WAVES>(sgj, tile - a.prefix[j]);
        v4f acc[RB][F16_NB];
        double2 wold[F16_NB];
        F16V V;
        f16_prefetch(sgj, base_i, a.wave, wold, a.overwrite != 0);
        f16_setup<ARITH>(a.plan, sPh, sgj, inv_pw, base_i, a.dt, V);
        f16_contract_prio<RB>(sTile, a.KB, V, acc, a.sched.prio);
        f16_epilogue<RB, ARITH>(a.plan, sTile, a.KB, sPh, sgj, inv_pw, inv_hw, base_i, a.dt, acc, V,
                                wold, a.wave);
    }
}

// Batch, persistent over the (waveform, interval, tile) sequence.
template <int RB, int WAVES, bool ARITH>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(WAVES / 4, WAVES / 4)))
void batch_fact16_kernel(BatchDev B, BatchRunDev r, BatchPtrs p, const unsigned char *tiles,
                         long tile_bytes, int KB, F16Sched sched)
{
    constexpr int NT = 64 * WAVES;
    extern __shared__ __attribute__((aligned(16))) unsigned char sTile[];
    const int tid = threadIdx.x;
    double *sPh = reinterpret_cast<double *>(sTile + tile_bytes);
    const int total = r.wt_prefix[B.B];
    const int per = (total + (int)gridDim.x - 1) / (int)gridDim.x;
    const int t0 = (int)blockIdx.x * per;
    const int t1 = min(total, t0 + per);
    if (r.overwrite) {
        // each waveform's samples after its last interval: zero (its intervals cover [0, end))
        for (int wb = 0; wb < B.B; wb++) {
            const WaveDesc dw = B.desc[wb];
            const long end = dw.L > 1 ? r.segs[dw.seg_off + dw.L - 2].end : 0;
            double2 *w = p.wave + dw.wave_off / 2;
            for (long i = end + (long)blockIdx.x * NT + tid; i < dw.N; i += (long)gridDim.x * NT)
                w[i] = make_double2(0.0, 0.0);
        }
    }
    if (t0 >= t1) return;
    int b = upper_index(r.wt_prefix, B.B, t0);
    WaveDesc d = B.desc[b];
    const int *pref = r.tpref + d.tpref_off;
    int j = upper_index(pref, d.L - 1, t0 - r.wt_prefix[b]);
    FactPlanDev P = batch_wave_plan(B, d);
    bool stale = true;
    Segment sgj;
    double inv_pw = 1.0, inv_hw = 1.0;  // per interval, computed when the interval changes
    for (int tile = t0; tile < t1; tile++) {
        while (tile >= r.wt_prefix[b + 1]) {  // next waveform with tiles
            b++;
            d = B.desc[b];
            pref = r.tpref + d.tpref_off;
            P = batch_wave_plan(B, d);
            j = 0;
            stale = true;
        }
        const int lt = tile - r.wt_prefix[b];
        while (lt >= pref[j + 1]) {
            j++;
            stale = true;
        }
        if (stale) {
            sgj = r.segs[d.seg_off + j];
            inv_pw = 1.0 / sgj.pw;
            inv_hw = 1.0 / sgj.hw;
            __syncthreads();
            const uint4 *src = reinterpret_cast<const uint4 *>(tiles + (long)(d.seg_off + j) * tile_bytes);
            uint4 *dst = reinterpret_cast<uint4 *>(sTile);
            const int nv = (int)(tile_bytes >> 4);
            for (int i = tid; i < nv; i += NT) dst[i] = src[i];
            const long np = (long)(d.L - 1) * 3;
            if (tid < 24) sPh[tid] = p.pcoef[d.pc_off + (tid & 7) * np + (long)j * 3 + (tid >> 3)];
            __syncthreads();
            stale = false;
        }
        double2 *wave = p.wave + d.wave_off / 2;
        const int base_i = f16_base<WAVES>(sgj, lt - pref[j]);
        v4f acc[RB][F16_NB];
        double2 wold[F16_NB];
        F16V V;
        f16_prefetch(sgj, base_i, wave, wold, r.overwrite != 0);
        f16_setup<ARITH>(P, sPh, sgj, inv_pw, base_i, d.dt, V);
        f16_contract_prio<RB>(sTile, KB, V, acc, sched.prio);
        f16_epilogue<RB, ARITH>(P, sTile, KB, sPh, sgj, inv_pw, inv_hw, base_i, d.dt, acc, V, wold,
                                wave);
    }
}

// Batch, row-streamed variant (batch_fact16_kernel's work split).
template <int KB, int WAVES>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(WAVES / 4, WAVES / 4)))
void batch_fact16s_kernel(BatchDev B, BatchRunDev r, BatchPtrs p, const unsigned char *tiles,
                          long tile_bytes, F16Sched sched)
{
    constexpr int NT = 64 * WAVES;
    extern __shared__ __attribute__((aligned(16))) unsigned char sTile[];
    const int tid = threadIdx.x;
    const int RB = B.uni.RB;
    double *sPh = reinterpret_cast<double *>(sTile + tile_bytes);
    const int total = r.wt_prefix[B.B];
    const int per = (total + (int)gridDim.x - 1) / (int)gridDim.x;
    const int t0 = (int)blockIdx.x * per;
    const int t1 = min(total, t0 + per);
    if (r.overwrite) {
        for (int wb = 0; wb < B.B; wb++) {
            const WaveDesc dw = B.desc[wb];
            const long end = dw.L > 1 ? r.segs[dw.seg_off + dw.L - 2].end : 0;
            double2 *w = p.wave + dw.wave_off / 2;
            for (long i = end + (long)blockIdx.x * NT + tid; i < dw.N; i += (long)gridDim.x * NT)
                w[i] = make_double2(0.0, 0.0);
        }
    }
    if (t0 >= t1) return;
    int b = upper_index(r.wt_prefix, B.B, t0);
    WaveDesc d = B.desc[b];
    const int *pref = r.tpref + d.tpref_off;
    int j = upper_index(pref, d.L - 1, t0 - r.wt_prefix[b]);
    FactPlanDev P = batch_wave_plan(B, d);
    bool stale = true;
    Segment sgj;
    double inv_pw = 1.0, inv_hw = 1.0;
    for (int tile = t0; tile < t1; tile++) {
        while (tile >= r.wt_prefix[b + 1]) {
            b++;
            d = B.desc[b];
            pref = r.tpref + d.tpref_off;
            P = batch_wave_plan(B, d);
            j = 0;
            stale = true;
        }
        const int lt = tile - r.wt_prefix[b];
        while (lt >= pref[j + 1]) {
            j++;
            stale = true;
        }
        if (stale) {
            sgj = r.segs[d.seg_off + j];
            inv_pw = 1.0 / sgj.pw;
            inv_hw = 1.0 / sgj.hw;
            __syncthreads();
            const uint4 *src = reinterpret_cast<const uint4 *>(tiles + (long)(d.seg_off + j) * tile_bytes);
            uint4 *dst = reinterpret_cast<uint4 *>(sTile);
            const int nv = (int)(tile_bytes >> 4);
            for (int i = tid; i < nv; i += NT) dst[i] = src[i];
            const long np = (long)(d.L - 1) * 3;
            if (tid < 24) sPh[tid] = p.pcoef[d.pc_off + (tid & 7) * np + (long)j * 3 + (tid >> 3)];
            __syncthreads();
            stale = false;
        }
        f16s_tile<KB>(P, sTile, RB, sPh, sgj, inv_pw, inv_hw, f16_base<WAVES>(sgj, lt - pref[j]),
                      d.dt, p.wave + d.wave_off / 2, r.overwrite != 0, sched.prio);
    }
}

template <int CPT>
__global__ F16_PREP_BOUNDS(CPT) void batch_fact16_prep_whole_kernel(
    BatchDev B, BatchPtrs p, unsigned char *tiles, long tile_bytes, int KB, int RB)
{
    const int gi = blockIdx.x;
    const int b = upper_index(B.seg_prefix, B.B, gi);
    const WaveDesc d = B.desc[b];
    const int j = gi - B.seg_prefix[b];
    fact16_prep_interval<CPT>(batch_wave_plan(B, d), p.interp + d.interp_off,
                         reinterpret_cast<const double2 *>(p.ylm + d.ylm_off), p.t + d.t_off, d.L,
                         j, tiles + (long)gi * tile_bytes, KB, RB);
}

int fact16_tile(int variant) { return 16 * F16_NB * fact16_waves(variant); }

static int device_cus16()
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

template <int WAVES, bool ARITH>
static int launch_fact16_w(const Fact16Args &a, int RB, int blocks, size_t lds, hipStream_t stream)
{
    switch (RB) {
#define FEW_CASE16(X)                                                                            \
    case X:                                                                                      \
        FEW_LDS((fact16_kernel<X, WAVES, ARITH>), lds);                                         \
        hipLaunchKernelGGL((fact16_kernel<X, WAVES, ARITH>), dim3(blocks), dim3(64 * WAVES),    \
                           lds, stream, a);                                                      \
        break;
        FEW_CASE16(1) FEW_CASE16(2) FEW_CASE16(3) FEW_CASE16(4) FEW_CASE16(5) FEW_CASE16(6)
        FEW_CASE16(7) FEW_CASE16(8) FEW_CASE16(9) FEW_CASE16(10) FEW_CASE16(11) FEW_CASE16(12)
        FEW_CASE16(13) FEW_CASE16(14) FEW_CASE16(15) FEW_CASE16(16)
#undef FEW_CASE16
    default:
        set_error("mixed16 sum: %d row blocks unsupported", RB);
        return FEW_ERR_UNSUPPORTED;
    }
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

template <int WAVES>
static int launch_fact16s_w(const Fact16Args &a, int RB, int blocks, size_t lds, hipStream_t stream)
{
    switch (a.KB) {
#define FEW_CASE16S(X)                                                                           \
    case X:                                                                                      \
        FEW_LDS((fact16s_kernel<X, WAVES>), lds);                                               \
        hipLaunchKernelGGL((fact16s_kernel<X, WAVES>), dim3(blocks), dim3(64 * WAVES), lds,     \
                           stream, a, RB);                                                       \
        break;
        FEW_CASE16S(1) FEW_CASE16S(2) FEW_CASE16S(3) FEW_CASE16S(4) FEW_CASE16S(5)
        FEW_CASE16S(6) FEW_CASE16S(7) FEW_CASE16S(8)
#undef FEW_CASE16S
    default:
        set_error("mixed16 sum: %d column steps unsupported by the streamed variant", a.KB);
        return FEW_ERR_UNSUPPORTED;
    }
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

int launch_fact16_sum(const Fact16Args &a_in, int RB, int max_blocks, int variant, hipStream_t stream)
{
    if (max_blocks <= 0) return FEW_OK;
    Fact16Args a = a_in;
    a.sched = fact16_sched(variant);
    const size_t lds = (size_t)a.tile_bytes + 24 * sizeof(double);
    if (lds > 160 * 1024) {
        set_error("mixed16 sum: coefficient tile of %zu bytes exceeds LDS", lds);
        return FEW_ERR_UNSUPPORTED;
    }
    if (variant & ~F16_VARIANT_BITS) {
        set_error("mixed16 sum: unknown variant %d", variant);
        return FEW_ERR_ARG;
    }
    int blocks = device_cus16() * (lds <= 80 * 1024 ? 2 : 1) * fact16_grid_factor(variant);
    if (blocks > max_blocks) blocks = max_blocks;
    const bool arith = a.plan.rows_arith != 0;
    const int waves = fact16_waves(variant);
    if (fact16_streamed(variant) && arith && a.KB >= 1 && a.KB <= F16S_MAX_KB) {
        return waves == 4 ? launch_fact16s_w<4>(a, RB, blocks, lds, stream)
                          : launch_fact16s_w<8>(a, RB, blocks, lds, stream);
    }
    if (waves == 4) {
        return arith ? launch_fact16_w<4, true>(a, RB, blocks, lds, stream)
                     : launch_fact16_w<4, false>(a, RB, blocks, lds, stream);
    }
    return arith ? launch_fact16_w<8, true>(a, RB, blocks, lds, stream)
                 : launch_fact16_w<8, false>(a, RB, blocks, lds, stream);
}

template <int WAVES, bool ARITH>
static int launch_batch_fact16_w(const BatchDev &B, const BatchRunDev &r, const BatchPtrs &p,
                                 const unsigned char *tiles, long tile_bytes, int KB,
                                 int blocks, size_t lds, F16Sched sched, hipStream_t stream)
{
    switch (B.uni.RB) {
#define FEW_CASEB(X)                                                                             \
    case X:                                                                                      \
        FEW_LDS((batch_fact16_kernel<X, WAVES, ARITH>), lds);                                   \
        hipLaunchKernelGGL((batch_fact16_kernel<X, WAVES, ARITH>), dim3(blocks),                \
                           dim3(64 * WAVES), lds, stream, B, r, p, tiles, tile_bytes, KB, sched); \
        break;
        FEW_CASEB(1) FEW_CASEB(2) FEW_CASEB(3) FEW_CASEB(4) FEW_CASEB(5) FEW_CASEB(6)
        FEW_CASEB(7) FEW_CASEB(8) FEW_CASEB(9) FEW_CASEB(10) FEW_CASEB(11) FEW_CASEB(12)
        FEW_CASEB(13) FEW_CASEB(14) FEW_CASEB(15) FEW_CASEB(16)
#undef FEW_CASEB
    default:
        set_error("batch mixed16 sum: %d row blocks unsupported", B.uni.RB);
        return FEW_ERR_UNSUPPORTED;
    }
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

template <int WAVES>
static int launch_batch_fact16s_w(const BatchDev &B, const BatchRunDev &r, const BatchPtrs &p,
                                  const unsigned char *tiles, long tile_bytes, int KB, int blocks,
                                  size_t lds, F16Sched sched, hipStream_t stream)
{
    switch (KB) {
#define FEW_CASEBS(X)                                                                            \
    case X:                                                                                      \
        FEW_LDS((batch_fact16s_kernel<X, WAVES>), lds);                                         \
        hipLaunchKernelGGL((batch_fact16s_kernel<X, WAVES>), dim3(blocks), dim3(64 * WAVES),    \
                           lds, stream, B, r, p, tiles, tile_bytes, sched);                      \
        break;
        FEW_CASEBS(1) FEW_CASEBS(2) FEW_CASEBS(3) FEW_CASEBS(4) FEW_CASEBS(5) FEW_CASEBS(6)
        FEW_CASEBS(7) FEW_CASEBS(8)
#undef FEW_CASEBS
    default:
        set_error("batch mixed16 sum: %d column steps unsupported by the streamed variant", KB);
        return FEW_ERR_UNSUPPORTED;
    }
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

int launch_batch_fact16(const BatchDev &B, const BatchRunDev &r, const BatchPtrs &p,
                        unsigned char *tiles, long tile_bytes, int KB, int variant,
                        int stage, hipStream_t stream)
{
    if (stage & 1) {
        if (B.total_intervals > 0) {
            const dim3 grid(B.total_intervals), block(F16_PREP_THREADS);
            const int RB = B.uni.RB;
            switch (f16_prep_cpt(KB, RB)) {
#define F16_BATCH_PREP(C)                                                                        \
    hipLaunchKernelGGL(batch_fact16_prep_whole_kernel<C>, grid, block, 0, stream, B, p, tiles,   \
                       tile_bytes, KB, RB)
            case 1: F16_BATCH_PREP(1); break;
            case 2: F16_BATCH_PREP(2); break;
            case 3: F16_BATCH_PREP(3); break;
            case 4: F16_BATCH_PREP(4); break;
            default: F16_BATCH_PREP(0); break;
#undef F16_BATCH_PREP
            }
            FEW_HIP(hipGetLastError());
        }
    }
    if (!(stage & 2)) return FEW_OK;
    const size_t lds = (size_t)tile_bytes + 24 * sizeof(double);
    if (lds > 160 * 1024) {
        set_error("batch mixed16 sum: coefficient tile of %zu bytes exceeds LDS", lds);
        return FEW_ERR_UNSUPPORTED;
    }
    if (variant & ~F16_VARIANT_BITS) {
        set_error("batch mixed16 sum: unknown variant %d", variant);
        return FEW_ERR_ARG;
    }
    BatchRunDev rs = r;
    rs.overwrite = (stage & 4) ? 1 : 0;
    const int blocks = device_cus16() * (lds <= 80 * 1024 ? 2 : 1) * fact16_grid_factor(variant);
    const bool arith = B.uni.rows_arith != 0;
    const F16Sched sched = fact16_sched(variant);
    const int waves = fact16_waves(variant);
    if (fact16_streamed(variant) && arith && KB >= 1 && KB <= F16S_MAX_KB)
        return waves == 4 ? launch_batch_fact16s_w<4>(B, rs, p, tiles, tile_bytes, KB, blocks, lds, sched, stream)
                          : launch_batch_fact16s_w<8>(B, rs, p, tiles, tile_bytes, KB, blocks, lds, sched, stream);
    if (waves == 4)
        return arith ? launch_batch_fact16_w<4, true>(B, rs, p, tiles, tile_bytes, KB, blocks, lds, sched, stream)
                     : launch_batch_fact16_w<4, false>(B, rs, p, tiles, tile_bytes, KB, blocks, lds, sched, stream);
    return arith ? launch_batch_fact16_w<8, true>(B, rs, p, tiles, tile_bytes, KB, blocks, lds, sched, stream)
                 : launch_batch_fact16_w<8, false>(B, rs, p, tiles, tile_bytes, KB, blocks, lds, sched, stream);
}

}  // namespace few

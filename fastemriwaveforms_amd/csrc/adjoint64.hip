// Factorised adjoint of the mode sum in FP64 on the matrix cores.
//
// Same algebra as the split-FP16 adjoint (adjoint16.hip): with u_r =
// exp(-i(m Phi_phi + k Phi_theta)), v_n = exp(-i n Phi_r) and sigma =
// (t - t_j)/h_j on knot interval j, every quantity the backward needs is a
// moment
//     M_j[r, n, s] = sum_{i in j} sigma_i^s conj(g_i) u_r(t_i) v_n(t_i),  s = 0..6,
// a (rows (r, s)) x (samples) x (columns n) contraction per interval.  Here
// it runs on v_mfma_f64_16x16x4f64 with FP64 operands and FP64 accumulation,
// so the moments carry FP64 rounding only.  That matters for the phase
// gradients: they are the small remainder of sums whose terms cancel by a
// factor of up to ~1e4 (a mode's own contribution to conj(g) e_q is constant
// in time when g resembles h), which FP32 accumulation cannot resolve.
//
// One workgroup owns one chunk of <= A64_CHUNK samples of one interval.  Per
// 16-sample step the workgroup writes, one job per thread, the A tile
// k-major (row 2R / 2R+1 of complex row R = 7 r + s, a = sigma^s conj(g) u_r,
// one job per (sample, r)) and the v table (one job per (sample, 8 columns):
// one sincos, then the z = exp(-i Phi_r) recurrence), then each wave runs the
// 8 MFMAs of a step for each of its 16 x 16 output blocks (column block cb,
// every NH-th row block).  The chunk's moments are written in FP64 and added
// per interval in a fixed order (adjoint64_reduce_kernel); the per-mode finish
// is the split-FP16 adjoint's (adjoint16_finish_kernel).
//
// Operand layout of v_mfma_f64_16x16x4f64 (checked on gfx950): lane l holds
// A[l & 15][l >> 4] and B[l >> 4][l & 15]; element i of its result is
// D[(l >> 4) + 4 i][l & 15].
#include <math.h>

#include "adjoint.hpp"
#include "batch.hpp"
#include "few_math.hpp"
#include "modesum_direct.hpp"
#include "modesum_fact.hpp"

namespace few {

typedef double v4d __attribute__((ext_vector_type(4)));

#define A64_STEP 16      // samples per step (K = 32 real per step: 8 MFMAs)
#define A64_SUPER 256    // samples whose phases are staged at a time
#define A64_VG 8         // v-table columns per job
#define INV_2PI_64 0.15915494309189533577

struct A64Layout {
    size_t a, b, pp, pt, pr, sig, cg, z, ph, total;
};

static __host__ __device__ inline size_t a64_align(size_t x) { return (x + 15) & ~(size_t)15; }

__host__ __device__ inline A64Layout a64_layout(int RBM, int NBC)
{
    A64Layout L;
    size_t o = 0;
    L.a = o;   o = a64_align(o + sizeof(double) * 2 * A64_STEP * (size_t)(RBM * 16 + 1));  // [k][row]
    L.b = o;   o = a64_align(o + sizeof(double) * 2 * A64_STEP * (size_t)(NBC * 16 + 1));  // [k][col]
    L.pp = o;  o = a64_align(o + sizeof(double) * A64_SUPER);
    L.pt = o;  o = a64_align(o + sizeof(double) * A64_SUPER);
    L.pr = o;  o = a64_align(o + sizeof(double) * A64_SUPER);
    L.sig = o; o = a64_align(o + sizeof(double) * A64_SUPER);
    L.cg = o;  o = a64_align(o + sizeof(double) * 2 * A64_SUPER);
    L.z = o;   o = a64_align(o + sizeof(double) * 2 * A64_SUPER);
    L.ph = o;  o = a64_align(o + sizeof(double) * 24);
    L.total = o;
    return L;
}

size_t adjoint64_lds_bytes(int RBM, int NBC) { return a64_layout(RBM, NBC).total; }

template <int RBM, int NH>
__global__ __launch_bounds__(512) void adjoint64_kernel(Adj16Args a)
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
    const int NRP = RBM * 16;       // padded real rows
    const int NCP = NBC * 16;       // padded columns
    // LDS strides of the k-rows: one double past the rows (columns), so the
    // 16 samples of a job group, which write one k-row each, fall on
    // different banks (at a multiple of 128 bytes all 16 hit one bank)
    const int SA = NRP + 1, SB = NCP + 1;
    const A64Layout Lay = a64_layout(RBM, NBC);
    double *sA = reinterpret_cast<double *>(smem + Lay.a);    // [k][row], k = 2 i + (re, im)
    double *sB = reinterpret_cast<double *>(smem + Lay.b);    // [k][col]
    double *sPp = reinterpret_cast<double *>(smem + Lay.pp);
    double *sPt = reinterpret_cast<double *>(smem + Lay.pt);
    double *sPr = reinterpret_cast<double *>(smem + Lay.pr);
    double *sSig = reinterpret_cast<double *>(smem + Lay.sig);
    double2 *sCg = reinterpret_cast<double2 *>(smem + Lay.cg);
    double2 *sZ = reinterpret_cast<double2 *>(smem + Lay.z);
    double *sPh = reinterpret_cast<double *>(smem + Lay.ph);

    const long np = (long)a.nseg * 3;
    if (tid < 24) sPh[tid] = a.pcoef[(tid & 7) * np + (long)j * 3 + (tid >> 3)];
    // padding rows of the A tile (real rows >= 14 Nr) stay zero for the whole chunk
    const int rows_used = 2 * A16_MOM * Nr;
    const int pad = NRP - rows_used;
    if (pad > 0) {
        for (int e = tid; e < 2 * A64_STEP * pad; e += NT) {
            const int k = e / pad, rr = e - k * pad;
            sA[k * SA + rows_used + rr] = 0.0;
        }
    }
    const double inv_pw = 1.0 / sg.pw;
    const double inv_hw = 1.0 / sg.hw;
    v4d acc[RBH];
#pragma unroll
    for (int k = 0; k < RBH; k++) acc[k] = v4d{0.0, 0.0, 0.0, 0.0};

    // job slots: (sample, row) jobs, then (sample, column group) jobs, each
    // kind starting on a wave boundary so no wave mixes the two
    const int u_slots = ((Nr * A64_STEP + 63) >> 6) << 6;
    const int v_jobs = (NCP / A64_VG) * A64_STEP;
    const int n_slots = u_slots + (((v_jobs + 63) >> 6) << 6);

    for (int sbase = lo; sbase < hi; sbase += A64_SUPER) {
        __syncthreads();
        for (int q = tid; q < A64_SUPER; q += NT) {
            const int i = sbase + q;
            const double t = a.dt * (double)i;
            const double s = (t - sg.pt) * inv_pw;
            sPp[q] = dense_phase(sPh, s, 1.0 - s) * INV_2PI_64;
            sPt[q] = P.has_k ? dense_phase(sPh + 8, s, 1.0 - s) * INV_2PI_64 : 0.0;
            const double pr = dense_phase(sPh + 16, s, 1.0 - s) * INV_2PI_64;
            sPr[q] = pr;
            double zs, zc;
            sincos_turns(reduce_turns(pr), zs, zc);
            sZ[q] = make_double2(zc, -zs);  // exp(-i Phi_r)
            sSig[q] = (t - sg.h_t) * inv_hw;
            double2 cg = make_double2(0.0, 0.0);
            if (i < hi) {
                const double2 gg = a.grad[i];
                cg = make_double2(gg.x, -gg.y);
            }
            sCg[q] = cg;
        }
        const int nsteps = (min(A64_SUPER, hi - sbase) + A64_STEP - 1) / A64_STEP;
        for (int st = 0; st < nsteps; st++) {
            const int i0 = st * A64_STEP;
            __syncthreads();
            for (int slot = tid; slot < n_slots; slot += NT) {
                if (slot < u_slots) {
                    // rows of complex row R = 7 r + s: a = sigma^s conj(g) u_r
                    const int r = slot / A64_STEP, ii = slot % A64_STEP;
                    if (r < Nr) {
                        const int2 mk = P.row_mk[r];
                        double sn, cs;
                        sincos_turns(reduce_turns((double)mk.x * sPp[i0 + ii] +
                                                  (double)mk.y * sPt[i0 + ii]),
                                     sn, cs);
                        const double ux = cs, uy = -sn;
                        const double2 cgv = sCg[i0 + ii];
                        double wx = cgv.x * ux - cgv.y * uy, wy = cgv.x * uy + cgv.y * ux;
                        const double sg1 = sSig[i0 + ii];
                        double *k0 = sA + (2 * ii) * SA, *k1 = k0 + SA;
#pragma unroll
                        for (int sm = 0; sm < A16_MOM; sm++) {
                            const int R2 = 2 * (A16_MOM * r + sm);
                            k0[R2] = wx;        // row 2R:   (Re, -Im)
                            k1[R2] = -wy;
                            k0[R2 + 1] = wy;    // row 2R+1: (Im, Re)
                            k1[R2 + 1] = wx;
                            wx *= sg1;
                            wy *= sg1;
                        }
                    }
                } else {
                    // v_n(t_i) for A64_VG consecutive columns
                    const int e = slot - u_slots;
                    if (e < v_jobs) {
                        const int ii = e % A64_STEP, grp = e / A64_STEP;
                        double sn, cs;
                        sincos_turns(reduce_turns((double)(P.n_lo + A64_VG * grp) * sPr[i0 + ii]),
                                     sn, cs);
                        double vr = cs, vi = -sn;
                        const double2 z = sZ[i0 + ii];
                        double *k0 = sB + (2 * ii) * SB, *k1 = k0 + SB;
#pragma unroll
                        for (int t2 = 0; t2 < A64_VG; t2++) {
                            const int col = A64_VG * grp + t2;
                            const bool in = col < P.Nc;
                            k0[col] = in ? vr : 0.0;
                            k1[col] = in ? vi : 0.0;
                            const double nr = vr * z.x - vi * z.y;
                            const double ni = vr * z.y + vi * z.x;
                            vr = nr;
                            vi = ni;
                        }
                    }
                }
            }
            // samples past the chunk end: their conj(g) is zero, so their A
            // columns are zero and they add nothing
            __syncthreads();
            double bf[2 * A64_STEP / 4];
#pragma unroll
            for (int kk = 0; kk < 2 * A64_STEP / 4; kk++) bf[kk] = sB[(4 * kk + g) * SB + 16 * cb + li];
#pragma unroll
            for (int k = 0; k < RBH; k++) {
                const int rb = part + NH * k;
                if (RBM % NH == 0 || rb < RBM) {
#pragma unroll
                    for (int kk = 0; kk < 2 * A64_STEP / 4; kk++) {
                        const double av = sA[(4 * kk + g) * SA + 16 * rb + li];
                        acc[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bf[kk], acc[k], 0, 0, 0);
                    }
                }
            }
        }
    }
    // chunk moments: D[16 rb + g + 4 v][16 cb + li]
    double *out = a.part64 + (long)c * NRP * NCP;
#pragma unroll
    for (int k = 0; k < RBH; k++) {
        const int rb = part + NH * k;
        if (RBM % NH == 0 || rb < RBM) {
#pragma unroll
            for (int v = 0; v < 4; v++)
                out[(long)(16 * rb + g + 4 * v) * NCP + 16 * cb + li] = acc[k][v];
        }
    }
}

// Folded variant (plans whose coefficient tiles fold column -n onto n).  For
// n >= 0 the moments of n and -n follow from four real sums of a = sigma^s
// conj(g) u_r (FP64) against v_n:
//     P = sum Re a Re v,  Q = sum Im a Im v,  X = sum Re a Im v,  S = sum Im a Re v,
//     M[r, n, s] = (P - Q) + i (X + S),   M[r, -n, s] = (P + Q) + i (S - X),
// a real contraction with samples as K (32 per step), rows (Re a, Im a) per
// (r, s) and columns (Re v_n, Im v_n) for n = 0 .. tcols - 1: half the MFMAs of
// the complex form, whose K holds (Re, Im) pairs of 16 samples over all n.
// The chunk tiles hold P, X (row 2R) and S, Q (row 2R + 1) at columns 2n, 2n + 1;
// adjoint64f_reduce_kernel forms M in the unfolded layout.
#define A64F_STEP 32     // samples per step (K = 32: 8 MFMAs)

template <int RBM, int NH>
__global__ __launch_bounds__(512) void adjoint64f_kernel(Adj16Args a)
{
    constexpr int RBH = (RBM + NH - 1) / NH;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const FactPlanDev &P = a.plan;
    const int NBC = a.NBC;  // column blocks of the folded tile: ceil(2 tcols / 16)
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
    const int NRP = RBM * 16;
    const int NCP = NBC * 16;
    const int SA = NRP + 1, SB = NCP + 1;  // k-row strides one double past 128-byte multiples
    const A64Layout Lay = a64_layout(RBM, NBC);  // 2 A64_STEP = A64F_STEP k-rows
    double *sA = reinterpret_cast<double *>(smem + Lay.a);    // [k][row], k = sample
    double *sB = reinterpret_cast<double *>(smem + Lay.b);    // [k][col]
    double *sPp = reinterpret_cast<double *>(smem + Lay.pp);
    double *sPt = reinterpret_cast<double *>(smem + Lay.pt);
    double *sPr = reinterpret_cast<double *>(smem + Lay.pr);
    double *sSig = reinterpret_cast<double *>(smem + Lay.sig);
    double2 *sCg = reinterpret_cast<double2 *>(smem + Lay.cg);
    double2 *sZ = reinterpret_cast<double2 *>(smem + Lay.z);
    double *sPh = reinterpret_cast<double *>(smem + Lay.ph);

    const long np = (long)a.nseg * 3;
    if (tid < 24) sPh[tid] = a.pcoef[(tid & 7) * np + (long)j * 3 + (tid >> 3)];
    const int rows_used = 2 * A16_MOM * Nr;
    const int pad = NRP - rows_used;
    if (pad > 0) {
        for (int e = tid; e < A64F_STEP * pad; e += NT) {
            const int k = e / pad, rr = e - k * pad;
            sA[k * SA + rows_used + rr] = 0.0;
        }
    }
    const double inv_pw = 1.0 / sg.pw;
    const double inv_hw = 1.0 / sg.hw;
    v4d acc[RBH];
#pragma unroll
    for (int k = 0; k < RBH; k++) acc[k] = v4d{0.0, 0.0, 0.0, 0.0};

    const int ncol = 2 * P.tcols;  // real columns in use
    const int u_slots = ((Nr * A64F_STEP + 63) >> 6) << 6;
    const int v_jobs = ((NCP / 2 + A64_VG - 1) / A64_VG) * A64F_STEP;
    const int n_slots = u_slots + (((v_jobs + 63) >> 6) << 6);

    for (int sbase = lo; sbase < hi; sbase += A64_SUPER) {
        __syncthreads();
        for (int q = tid; q < A64_SUPER; q += NT) {
            const int i = sbase + q;
            const double t = a.dt * (double)i;
            const double s = (t - sg.pt) * inv_pw;
            sPp[q] = dense_phase(sPh, s, 1.0 - s) * INV_2PI_64;
            sPt[q] = P.has_k ? dense_phase(sPh + 8, s, 1.0 - s) * INV_2PI_64 : 0.0;
            const double pr = dense_phase(sPh + 16, s, 1.0 - s) * INV_2PI_64;
            sPr[q] = pr;
            double zs, zc;
            sincos_turns(reduce_turns(pr), zs, zc);
            sZ[q] = make_double2(zc, -zs);  // exp(-i Phi_r)
            sSig[q] = (t - sg.h_t) * inv_hw;
            double2 cg = make_double2(0.0, 0.0);
            if (i < hi) {
                const double2 gg = a.grad[i];
                cg = make_double2(gg.x, -gg.y);
            }
            sCg[q] = cg;
        }
        const int nsteps = (min(A64_SUPER, hi - sbase) + A64F_STEP - 1) / A64F_STEP;
        for (int st = 0; st < nsteps; st++) {
            const int i0 = st * A64F_STEP;
            __syncthreads();
            for (int slot = tid; slot < n_slots; slot += NT) {
                if (slot < u_slots) {
                    // rows 2R, 2R + 1 of R = 7 r + s: Re, Im of a = sigma^s conj(g) u_r
                    const int r = slot / A64F_STEP, ii = slot % A64F_STEP;
                    if (r < Nr) {
                        const int2 mk = P.row_mk[r];
                        double sn, cs;
                        sincos_turns(reduce_turns((double)mk.x * sPp[i0 + ii] +
                                                  (double)mk.y * sPt[i0 + ii]),
                                     sn, cs);
                        const double ux = cs, uy = -sn;
                        const double2 cgv = sCg[i0 + ii];
                        double wx = cgv.x * ux - cgv.y * uy, wy = cgv.x * uy + cgv.y * ux;
                        const double sg1 = sSig[i0 + ii];
                        double *k0 = sA + ii * SA;
#pragma unroll
                        for (int sm = 0; sm < A16_MOM; sm++) {
                            const int R2 = 2 * (A16_MOM * r + sm);
                            k0[R2] = wx;
                            k0[R2 + 1] = wy;
                            wx *= sg1;
                            wy *= sg1;
                        }
                    }
                } else {
                    // v_n(t_i) for A64_VG consecutive n >= 0: columns 2n (Re), 2n + 1 (Im)
                    const int e = slot - u_slots;
                    if (e < v_jobs) {
                        const int ii = e % A64F_STEP, grp = e / A64F_STEP;
                        double sn, cs;
                        sincos_turns(reduce_turns((double)(A64_VG * grp) * sPr[i0 + ii]), sn, cs);
                        double vr = cs, vi = -sn;
                        const double2 z = sZ[i0 + ii];
                        double *k0 = sB + ii * SB;
#pragma unroll
                        for (int t2 = 0; t2 < A64_VG; t2++) {
                            const int col = 2 * (A64_VG * grp + t2);
                            if (col < NCP) {
                                const bool in = col < ncol;
                                k0[col] = in ? vr : 0.0;
                                k0[col + 1] = in ? vi : 0.0;
                            }
                            const double nr = vr * z.x - vi * z.y;
                            const double ni = vr * z.y + vi * z.x;
                            vr = nr;
                            vi = ni;
                        }
                    }
                }
            }
            __syncthreads();
            double bf[A64F_STEP / 4];
#pragma unroll
            for (int kk = 0; kk < A64F_STEP / 4; kk++) bf[kk] = sB[(4 * kk + g) * SB + 16 * cb + li];
#pragma unroll
            for (int k = 0; k < RBH; k++) {
                const int rb = part + NH * k;
                if (RBM % NH == 0 || rb < RBM) {
#pragma unroll
                    for (int kk = 0; kk < A64F_STEP / 4; kk++) {
                        const double av = sA[(4 * kk + g) * SA + 16 * rb + li];
                        acc[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bf[kk], acc[k], 0, 0, 0);
                    }
                }
            }
        }
    }
    double *out = a.part64 + (long)c * NRP * NCP;
#pragma unroll
    for (int k = 0; k < RBH; k++) {
        const int rb = part + NH * k;
        if (RBM % NH == 0 || rb < RBM) {
#pragma unroll
            for (int v = 0; v < 4; v++)
                out[(long)(16 * rb + g + 4 * v) * NCP + 16 * cb + li] = acc[k][v];
        }
    }
}

// M_j in the unfolded layout (row 2R + e, column n - n_lo) from the folded
// chunk tiles of interval j, each of the two sums it combines added over the
// chunks in order (FP64).
__global__ __launch_bounds__(256) void adjoint64f_reduce_kernel(const double *part, const int *prefix,
                                                                int nseg, int ncp_f, long tile_f,
                                                                int ncp, long tile, int Nc, int n_lo,
                                                                double *M)
{
    const int j = blockIdx.y;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= tile) return;
    const int row = (int)(e / ncp), col = (int)(e - (long)row * ncp);
    double out = 0.0;
    if (col < Nc) {
        const int n = n_lo + col, m = n < 0 ? -n : n;
        const long b = (long)(row & ~1) * ncp_f + 2 * m;  // P at b, X at b + 1, S / Q one row on
        // Re M: P - Q (n >= 0), P + Q (n < 0); Im M: X + S (n >= 0), S - X (n < 0)
        const long ix = (row & 1) ? b + ncp_f : b;
        const long iy = (row & 1) ? b + 1 : b + ncp_f + 1;
        const double sgn = ((row & 1) != 0) == (n >= 0) ? 1.0 : -1.0;
        double sx = 0.0, sy = 0.0;
        for (int c = prefix[j]; c < prefix[j + 1]; c++) {
            sx += part[(long)c * tile_f + ix];
            sy += part[(long)c * tile_f + iy];
        }
        out = sx + sgn * sy;
    }
    M[(long)j * tile + e] = out;
}

// M_j = sum over the chunks of interval j (FP64, chunk order); zero for
// intervals without samples.
__global__ __launch_bounds__(256) void adjoint64_reduce_kernel(const double *part, const int *prefix,
                                                               int nseg, long tile, double *M)
{
    const int j = blockIdx.y;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= tile) return;
    double acc = 0.0;
    for (int c = prefix[j]; c < prefix[j + 1]; c++) acc += part[(long)c * tile + e];
    M[(long)j * tile + e] = acc;
}

static inline int a64_nh(int NBC) { return NBC <= 4 ? 2 : 1; }

int launch_adjoint64(const Adj16Args &a, int max_chunks, hipStream_t stream)
{
    if (max_chunks <= 0) return FEW_OK;
    const size_t lds = adjoint64_lds_bytes(a.RBM, a.NBC);
    if (lds > 160 * 1024 || a.NBC < 1 || a.NBC > 8) {
        set_error("FP64 factorised adjoint: tile of %zu LDS bytes / %d column blocks unsupported",
                  lds, a.NBC);
        return FEW_ERR_UNSUPPORTED;
    }
    const int nh = a64_nh(a.NBC);
    const dim3 block(64 * a.NBC * nh);
    const bool fold = a.plan.fold != 0;
    switch (a.RBM) {
#define A64_LAUNCH(K)                                                                           \
    do {                                                                                        \
        FEW_LDS(K, lds);                                                                        \
        hipLaunchKernelGGL(K, dim3(max_chunks), block, lds, stream, a);                         \
    } while (0)
#define A64_CASE(X)                                                                             \
    case X:                                                                                     \
        if (fold) {                                                                             \
            if (nh == 2) A64_LAUNCH((adjoint64f_kernel<X, 2>));                                 \
            else A64_LAUNCH((adjoint64f_kernel<X, 1>));                                         \
        } else {                                                                                \
            if (nh == 2) A64_LAUNCH((adjoint64_kernel<X, 2>));                                  \
            else A64_LAUNCH((adjoint64_kernel<X, 1>));                                          \
        }                                                                                       \
        break;
        A64_CASE(1) A64_CASE(2) A64_CASE(3) A64_CASE(4) A64_CASE(5) A64_CASE(6) A64_CASE(7)
        A64_CASE(8) A64_CASE(9) A64_CASE(10) A64_CASE(11) A64_CASE(12) A64_CASE(13) A64_CASE(14)
        A64_CASE(15) A64_CASE(16) A64_CASE(17) A64_CASE(18) A64_CASE(19) A64_CASE(20) A64_CASE(21)
        A64_CASE(22) A64_CASE(23) A64_CASE(24)
#undef A64_CASE
#undef A64_LAUNCH
    default:
        set_error("FP64 factorised adjoint: %d row blocks unsupported", a.RBM);
        return FEW_ERR_UNSUPPORTED;
    }
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

int launch_adjoint64_reduce(const Adj16Args &a, long tile, int ncol, double *M, hipStream_t stream)
{
    if (a.nseg <= 0) return FEW_OK;
    const dim3 grid((unsigned)((tile + 255) / 256), a.nseg);
    if (a.plan.fold) {
        const int ncp_f = a.NBC * 16;
        const long tile_f = (long)a.RBM * 16 * ncp_f;
        hipLaunchKernelGGL(adjoint64f_reduce_kernel, grid, dim3(256), 0, stream, a.part64,
                           a.prefix, a.nseg, ncp_f, tile_f, ncol, tile, a.plan.Nc, a.plan.n_lo, M);
    } else {
        hipLaunchKernelGGL(adjoint64_reduce_kernel, grid, dim3(256), 0, stream, a.part64, a.prefix,
                           a.nseg, tile, M);
    }
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

}  // namespace few

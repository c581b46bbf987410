// Factorised mode sum: shared declarations.
//
// The phase of mode q is m_q Phi_phi + k_q Phi_theta + n_q Phi_r, so
//   exp(-i phase_q) = u_r(t) * v_n(t),
//   u_r = exp(-i (m Phi_phi + k Phi_theta)),  v_n = exp(-i n Phi_r).
// Within knot interval j the amplitude of mode q is a cubic in
// s = (t - t_j) / (t_{j+1} - t_j).  Grouping modes (and the conjugate
// -m partners of interpolate.cu:665-677) by row r = signed (m, k) and
// column n gives
//   h(t) = sum_p s^p sum_r u_r(t) sum_n C_p[r, n] v_n(t),
// where C_p[r, n] (complex) collects Y+ c_{p,q} h_j^p and Y- conj(c_{p,q}) h_j^p.
// For a tile of samples inside one interval the inner sum over n is a dense
// GEMM (rows (r, p) x columns n) x (n x samples), run on MFMA; the outer sums
// are an elementwise epilogue.
#pragma once

#include <stdint.h>

#include <vector>

#include "few_common.hpp"

namespace few {

// Host-built description of a mode set (rows / columns / gather lists).
struct FactPlanHost {
    int K = 0;              // modes
    int Nr = 0;             // rows: distinct signed (m, k)
    int Nc = 0;             // columns: n_lo .. n_lo + Nc - 1
    int n_lo = 0;
    int RB = 0;             // 16-row MFMA blocks: ceil(8 * Nr / 16) (4 powers x re/im)
    int KS = 0;             // 4-deep k-steps: ceil(2 * tcols / 4)
    int CST = 0;            // LDS row stride in pairs (== 2 mod 16): >= 2 KS, 4 KS folded
    int rows_arith = 0;     // rows are m0, m0+1, ... with one k: phasors by recurrence
    int m0 = 0, k0 = 0;
    int has_k = 0;          // some mode has k != 0 (Phi_theta enters the phase)
    // Columns of the coefficient tiles.  v_{-n} = conj(v_n), so with fold the
    // coefficients of columns n and -n share tile column n (n = 0 .. tcols - 1)
    // as a real 2 x 2 block (fact_cell); otherwise tile column c is n = tn0 + c.
    // Folding is taken when the n range reaches below zero and it shortens the tiles.
    int fold = 0, tcols = 0, tn0 = 0;
    std::vector<int> row_m, row_k;        // (Nr,)
    std::vector<int> cell_start;          // (Nr*Nc + 1,)
    std::vector<int> entries;             // q*2 + conj, grouped by cell, q ascending
    std::vector<int> mode_cell;           // (2K,): cell of mode q, cell of its conjugate term (-1: none)
    bool ok = false;
};

// Device-side view of an uploaded plan (all pointers into one allocation).
struct FactPlanDev {
    int K, Nr, Nc, n_lo, RB, KS, CST;
    int rows_arith, m0, k0, has_k;
    int fold, tcols, tn0;
    const int2 *row_mk;      // (Nr,)
    const int *cell_start;   // (Nr*Nc + 1,)
    const int *entries;
    const int *mode_cell;    // (2K,)
};

struct FactPlan {
    FactPlanHost host;
    FactPlanDev dev;
    void *dev_buf = nullptr;
    size_t dev_bytes = 0;
    int device = -1;
};

// Largest supported sizes of the MFMA kernels (others fall back to the direct sum).
#define FACT_MAX_RB 16
#define FACT_MAX_NC 512

bool build_fact_plan(const int *m, const int *k, const int *n, int K, FactPlanHost &p);

// Coefficient tiles of the FP32 / FP64 sums: per interval, (RB*8 rows) x
// (CST pairs).  Unfolded: C[R, c] at (R * CST + c) * 2 + {re, im}.  Folded:
// the 2 x 2 block of tile column c, entry (row 2R + e, column 2c + f), at
// (R * CST + f * 2 KS + c) * 2 + e (the f = 0 and f = 1 halves of a row side
// by side, each 2 KS pairs wide).
inline size_t fact_coef_floats_per_interval(const FactPlanHost &p)
{
    return (size_t)p.RB * 8 * p.CST * 2;
}

// Adds C_p[r, col] (the entries of one cell, in list order) for the four
// powers: sum over the cell's modes q of Y c_{p,q} (conjugate entries:
// Y- conj(c_{p,q})), from the amplitude coefficients interp (4, L, 2K).
__device__ __forceinline__ void fact_gather(const FactPlanDev &P, const double *interp,
                                            const double2 *ylm, int L, int j, int cell,
                                            double (&cre)[4], double (&cim)[4])
{
    const long K2 = 2L * P.K;
    for (int e = P.cell_start[cell]; e < P.cell_start[cell + 1]; e++) {
        const int ent = P.entries[e];
        const int qm = ent >> 1;
        const bool cj = ent & 1;
        const double2 Y = ylm[cj ? P.K + qm : qm];
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const long base = ((long)p * L + j) * K2;
            const double cr = interp[base + qm];
            const double ci = cj ? -interp[base + P.K + qm] : interp[base + P.K + qm];
            cre[p] += cr * Y.x - ci * Y.y;
            cim[p] += cr * Y.y + ci * Y.x;
        }
    }
}

// The real 2 x 2 block of complex row r, tile column kc and each power p,
// scaled by h^p: x[p] = {(2R, 2kc), (2R, 2kc + 1), (2R + 1, 2kc), (2R + 1, 2kc + 1)}
// of R = 4 r + p, against the B rows (Re v_n, Im v_n) of tile column kc.
// With C = C_p[r, n] and D = C_p[r, -n] (D = 0 unless the tiles fold and n > 0),
//   C v_n + D conj(v_n) = (Cr + Dr) Re v - (Ci - Di) Im v
//                       + i [(Ci + Di) Re v + (Cr - Dr) Im v].
// Padding rows and columns are zero.
__device__ __forceinline__ void fact_cell(const FactPlanDev &P, const double *interp,
                                          const double2 *ylm, int L, int j, double hw, int r,
                                          int kc, double (&x)[4][4])
{
    double cre[4] = {0.0, 0.0, 0.0, 0.0}, cim[4] = {0.0, 0.0, 0.0, 0.0};
    double dre[4] = {0.0, 0.0, 0.0, 0.0}, dim[4] = {0.0, 0.0, 0.0, 0.0};
    if (r < P.Nr && kc < P.tcols) {
        const int n = P.tn0 + kc;
        const int col = n - P.n_lo;
        if (col < P.Nc) fact_gather(P, interp, ylm, L, j, r * P.Nc + col, cre, cim);
        const int colm = -n - P.n_lo;
        if (P.fold && n > 0 && colm >= 0) fact_gather(P, interp, ylm, L, j, r * P.Nc + colm, dre, dim);
    }
    double hp = 1.0;
#pragma unroll
    for (int p = 0; p < 4; p++) {
        x[p][0] = (cre[p] + dre[p]) * hp;
        x[p][1] = (dim[p] - cim[p]) * hp;
        x[p][2] = (cim[p] + dim[p]) * hp;
        x[p][3] = (cre[p] - dre[p]) * hp;
        hp *= hw;
    }
}

struct FactArgs {
    double2 *wave;
    const double *pcoef;          // (8, L-1, 3)
    const Segment *segs;
    const int *prefix;
    int nseg;
    double dt;
    const float *C;               // (L-1) x RB*8 x CST x 2
    long c_stride;                // floats per interval (doubles for C64)
    FactPlanDev plan;
    const double *C64 = nullptr;  // the FP64 tiles of the fact64 sum (same layout)
    int overwrite = 0;            // fact64: write h instead of adding; zero the tail
    int out_len = 0;
};

struct FactPrepArgs {
    const double *interp;         // (4, L, 2K)
    const double2 *ylm;           // (2K,)
    const double *h_t;            // (L,) amplitude knots (device)
    int L;
    float *C;
    long c_stride;
    FactPlanDev plan;
    double *C64 = nullptr;        // when set: FP64 tiles instead of C
};

int launch_fact_prep(const FactPrepArgs &a, hipStream_t stream);

// ---- FP64 factorised sum (modesum_fact.hip, precision FEW_PRECISION_FACT64):
// the same tiles in FP64 and v_mfma_f64_16x16x4f64, FP64 phasors and epilogue.
// variant 0: 8 waves x 16 samples per workgroup tile, 4: 4 waves x 32.
int fact64_tile(int variant);
int launch_fact64_sum(const FactArgs &a, const FactPlanHost &p, int max_blocks, int variant,
                      hipStream_t stream);

// ---- FP16 split-operand variant (modesum_fact16.hip)
// Tile of one interval: two FP16 planes (hi, lo), each KB x RB blocks of
// 16 real rows x 32 halfs (block (kb, rb) at (kb * RB + rb) KiB), then 4
// floats of inverse scales (one per power p), padded to 256 bytes.  The four 16-byte
// chunks of a row are XOR-swizzled by (row >> 1) & 3, which puts each group of
// eight lanes of a fragment read on distinct LDS banks without padding.
inline int fact16_kb(const FactPlanHost &p) { return (2 * p.tcols + 31) / 32; }
inline size_t fact16_tile_bytes(const FactPlanHost &p)
{
    const size_t b = 2 * (size_t)fact16_kb(p) * p.RB * 1024 + 4 * sizeof(float);
    return (b + 255) & ~(size_t)255;
}

struct Fact16PrepArgs {
    const double *interp;
    const double2 *ylm;
    const double *h_t;
    int L;
    unsigned char *tiles;
    long tile_bytes;
    int KB, RB;
    FactPlanDev plan;
};

// Wave priority of the mixed16 sum kernels (modesum_fact16.hip): with prio
// a wave raises its priority over its MFMA contraction, so the two waves
// sharing a SIMD interleave one's MFMAs with the other's VALU work.
struct F16Sched {
    int prio;
};
// variant of the split-FP16 sum, a bit set:
//   4  -- 4 waves per workgroup (one per SIMD) instead of 8;
//   8  -- no raised wave priority over the contraction;
//   16 -- the row-streamed contraction (taken for arithmetic rows and at most
//         F16S_MAX_KB column steps; otherwise the regular one runs);
//   32 -- twice as many persistent workgroups as fit at once (shorter
//         contiguous tile ranges, balanced by the hardware scheduler).
constexpr int F16_VARIANT_BITS = 4 | 8 | 16 | 32;
inline F16Sched fact16_sched(int variant) { return F16Sched{(variant & 8) ? 0 : 1}; }
inline int fact16_waves(int variant) { return (variant & 4) ? 4 : 8; }
inline bool fact16_streamed(int variant) { return (variant & 16) != 0; }
inline int fact16_grid_factor(int variant) { return (variant & 32) ? 2 : 1; }
constexpr int F16S_MAX_KB = 8;

struct Fact16Args {
    double2 *wave;
    const double *pcoef;
    const Segment *segs;
    const int *prefix;
    int nseg;
    double dt;
    const unsigned char *tiles;
    long tile_bytes;
    int KB;
    FactPlanDev plan;
    F16Sched sched;
    int overwrite = 0;  // write h instead of adding it; zero the samples after the last interval
    int out_len = 0;
};

int launch_fact16_prep(const Fact16PrepArgs &a, hipStream_t stream);
int fact16_tile(int variant);
int launch_fact16_sum(const Fact16Args &a, int RB, int max_blocks, int variant, hipStream_t stream);
int fact_tile(const FactPlanHost &p, int variant);
int launch_fact_sum(const FactArgs &a, const FactPlanHost &p, int max_blocks, int variant,
                    hipStream_t stream);

}  // namespace few

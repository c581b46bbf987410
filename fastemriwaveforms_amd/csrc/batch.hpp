// Batched mode summation: many waveforms with their own durations, knot
// counts and mode subsets in one launch per stage.
//
// Inputs and outputs are packed back to back (per-waveform offsets below).
// Every waveform's modes are mapped onto one union (row = signed (m, k),
// column = n) grid so all waveforms share the coefficient-tile shape of the
// factorised kernels; a cell a waveform does not populate is zero.
#pragma once

#include "few_common.hpp"
#include "modesum_fact.hpp"

namespace few {

struct WaveDesc {
    int L, K, N;          // knots, modes, output samples
    int seg_off;          // first global interval index: sum of (L - 1) before
    int tpref_off;        // first entry of this waveform's tile prefix (L entries)
    int sys_off;          // first spline system: sum of (2K + 3) before
    int mode_off;         // first mode: sum of K before
    int cell_off;         // first cell_start entry: b * (Nr*Nc + 1)
    int ent_off;          // first gather entry
    int pad_;
    double dt;
    long t_off;           // knot times (L)
    long teuk_off;        // knot amplitudes, complex (L, K) -> doubles
    long pk_off;          // phase knots (3, L)
    long ylm_off;         // Y+ then Y-, complex (2K) -> doubles
    long interp_off;      // amplitude coefficients (4, L, 2K)
    long pc_off;          // dense-output phase coefficients (8, L-1, 3)
    long wave_off;        // output, complex (N) -> doubles
    long sc_off;          // spline scratch (batch_spline_scratch(L, K) doubles)
};

// Spline scratch of one waveform: its knot matrix eliminated once (dgtsv's
// piv | fact | d | du | dl, 5L doubles) and that elimination's INFO, padded
// to 64 bytes, then four (L, 2K+3) planes.  Every system of the waveform
// shares the matrix; with INFO = 0 the systems sweep their right-hand sides
// in the first plane, otherwise each system runs dgtsv on its own in all four
// (LAPACK's early return depends on the right-hand side).
__host__ __device__ inline long batch_spline_head(long L) { return (5 * L + 1 + 7) & ~7L; }
__host__ __device__ inline long batch_spline_scratch(long L, long K)
{
    return batch_spline_head(L) + 4 * L * (2 * K + 3);
}

// Device tables of a batch plan (one allocation).
struct BatchDev {
    int B;
    int total_intervals;
    long total_systems;
    const WaveDesc *desc;     // (B,)
    const long *sys_prefix;   // (B+1,) spline systems
    const int *seg_prefix;    // (B+1,) intervals
    const int *cell_start;    // (B * (Nr*Nc + 1),)
    const int *entries;
    const int *m, *k, *n;     // concatenated mode indices (sum K)
    FactPlanDev uni;          // union grid (row tables); K, cell_start, entries unused
};

// Mutable per-run arrays (in the caller's workspace).
struct BatchRunDev {
    Segment *segs;       // (total_intervals,)
    int *tpref;          // (sum L,) per-waveform exclusive tile prefix + total
    int *wt_prefix;      // (B+1,) tiles before each waveform
    int overwrite;       // sum writes every output sample (0 after each waveform's last interval)
};

struct BatchPtrs {
    const double *t, *teuk, *pk, *ylm;
    double *interp, *pcoef, *scratch;
    double2 *wave;
};

// A waveform's view of the union plan (its own K and gather lists).
__host__ __device__ inline FactPlanDev batch_wave_plan(const BatchDev &B, const WaveDesc &d)
{
    FactPlanDev P = B.uni;
    P.K = d.K;
    P.cell_start = B.cell_start + d.cell_off;
    P.entries = B.entries + d.ent_off;
    return P;
}

// Binary search: largest i in [0, n) with prefix[i] <= x (prefix[0] <= x < prefix[n]).
template <typename T>
__host__ __device__ inline int upper_index(const T *prefix, int n, T x)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

int launch_batch_splines(const BatchDev &B, const BatchPtrs &p, hipStream_t stream);
int launch_batch_plan(const BatchDev &B, const BatchRunDev &r, const BatchPtrs &p, int tile,
                      hipStream_t stream);
int launch_batch_fact16(const BatchDev &B, const BatchRunDev &r, const BatchPtrs &p,
                        unsigned char *tiles, long tile_bytes, int KB, int variant,
                        int stage, hipStream_t stream);
int launch_batch_direct_fp64(const BatchDev &B, const BatchRunDev &r, const BatchPtrs &p,
                             int max_blocks, hipStream_t stream);

}  // namespace few

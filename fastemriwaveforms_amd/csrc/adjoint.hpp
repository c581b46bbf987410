// Adjoint (reverse mode) of the interpolated mode sum and of the spline
// build, for a real loss L with upstream gradient g = dL/dh, i.e.
// L = sum_i (g_re h_re + g_im h_im) to first order.
//
// Mode sum (make_waveform, interpolate.cu:618-686).  For a sample of
// interval j with x = t - h_t[j], e = exp(-i phi), A the amplitude cubic:
//   W = conj(g) e Y+,  V = [m != 0] conj(g) conj(e) Y-
//   dL/dRe A = Re(W + V),  dL/dIm A = Im(V - W)
//   dL/dphi  = Im(A) dL/dRe A - Re(A) dL/dIm A
// and phi = m Phi_phi + k Phi_theta + n Phi_r.  The coefficients of
// interval j collect x^p times these over the interval's samples
// (segmented per-mode reductions); the phase cubics of interval j collect
// xp^p dL/dPhi with xp = t - phase_t[j].
//
// Spline build (interpolate.cu:284-325): linear in the knot values for
// fixed knots, so its adjoint is the transpose: final derivatives folded
// back, fill_coefficients transposed, one transposed tridiagonal solve
// A^T lambda = dL/dz, then fill_B transposed.
#pragma once

#include "few_common.hpp"
#include "modesum_fact.hpp"

namespace few {

struct AdjArgs {
    const double2 *grad;      // (out_len,) dL/dh
    const double *interp;     // (4, L, 2K) amplitude coefficients
    const double *pcoef;      // (8, L-1, 3) dense-output phase coefficients
    const int *m, *k, *n;     // (K,)
    const double2 *ylm;       // (2K,) Y+ then Y-
    const Segment *segs;      // (L-1,) from the planner, chunk = tile
    const int *prefix;        // (L,) exclusive chunk prefix
    int nseg, L, K;
    double dt;
    int chunk;                // samples per chunk
    int groups;               // mode groups (blockIdx.y)
    double *part_amp;         // (max_chunks, 4, 2K) per-chunk partial dL/dc
    double *part_phase;       // (max_chunks, groups, 12) per-chunk partial phase moments
};

struct AdjReduceArgs {
    const int *prefix;
    const double *part_amp;
    const double *part_phase;
    int nseg, L, K, groups;
    double *g_interp;         // (4, L, 2K) out
    double *g_phase;          // (4, L, 3) out
};

// One transposed spline system: gradient G(p, i) = g[p*gp + i*gi] of the
// (4, length) coefficients of one array -> dL/dy(i) = gy[i*gyi].
struct SplineAdjGroup {
    const double *g;   // coefficient gradient of system 0
    long g_sys, g_p, g_knot;
    double *gy;        // knot-value gradient of system 0
    long gy_sys, gy_knot;
    int nsys;
};

struct SplineAdjArgs {
    const double *t;          // (L,) knots shared by every system
    SplineAdjGroup grp[3];
    int ngroups;
    int L;
    long ns;                  // scratch columns (total systems)
    double *scratch;          // L * ns doubles: one right-hand side per system
    double *factor;           // 7 * L doubles: the shared elimination (tri_factor_transposed)
    int blk_prefix[4];        // first workgroup of each group (LDS path)
};

// Knot counts up to this run the LDS path: knots, elimination and two L x 32
// tiles of one workgroup's right-hand sides stay in LDS (8 L (8 + 64) bytes).
#define ADJ_SHARED_MAX_L 150
#define ADJ_PF 8               // prefetch distance (intervals) of the gradient loads

// Elimination of the transposed not-a-knot matrix of knots x (shared by all
// systems on those knots) into F (7 x L doubles); returns dgtsv's INFO.
__host__ __device__ int tri_factor_transposed(const double *x, int L, double *F);
// The transpose of one system's build with the shared elimination F;
// s holds L scratch doubles with stride ws.
__host__ __device__ void spline_adjoint_rhs(const double *x, int L, const double *F,
                                            const double *g, long gp, long gk, double *gy,
                                            long gyk, double *s, long ws);

// ---- mixed precision (adjoint16.hip): moments on split-FP16 MFMA
#define A16_MOM 7              // moments sigma^0 .. sigma^6
#define A16_CHUNK 1024         // samples per chunk (one workgroup)

struct Adj16Args {
    const double2 *grad;       // (out_len,) dL/dh
    const double *pcoef;       // (8, L-1, 3)
    const Segment *segs;       // planner output, tile = chunk
    const int *prefix;
    int nseg;
    double dt;
    int chunk;
    FactPlanDev plan;
    int RBM;                   // 16-row blocks of the moment tile: ceil(2 * 7 * Nr / 16)
    int NBC;                   // 16-column blocks (waves): ceil(Nc / 16); FP64 with folded
                               // tiles: ceil(2 tcols / 16)
    float *part;               // (max_chunks, RBM*16, NBC*16) chunk moments (split FP16)
    double *part64;            // (max_chunks, RBM*16, NBC*16) chunk moments (FP64 MFMA)
};

struct Adj16FinishArgs {
    FactPlanDev plan;
    double *M;                 // (L-1, RBM*16, NBC*16) interval moments (FP64)
    long tile;                 // RBM*16 * NBC*16
    int ncol;                  // NBC*16
    const double *h_t;         // (L,)
    int L;
    const double *interp;      // (4, L, 2K)
    const double2 *ylm;
    const int *m, *k, *n;
    double *g_interp;          // (4, L, 2K) out
    double *g_phase;           // (4, L, 3) out
    double *part_phase;        // (L, groups, 12)
};

size_t adjoint16_lds_bytes(int RBM, int NBC, int Nr);
#define A64_CHUNK 2048         // samples per chunk of the FP64 factorised adjoint
size_t adjoint64_lds_bytes(int RBM, int NBC);
int launch_adjoint64(const Adj16Args &a, int max_chunks, hipStream_t stream);
int launch_adjoint64_reduce(const Adj16Args &a, long tile, int ncol, double *M, hipStream_t stream);
int launch_adjoint16(const Adj16Args &a, int max_chunks, hipStream_t stream);
// reduce_chunks: add the FP32 chunk moments of the split-FP16 adjoint into f.M
// first (the FP64 factorised adjoint reduces its own chunks)
int launch_adjoint16_finish(const Adj16Args &a, const Adj16FinishArgs &f, hipStream_t stream,
                            bool reduce_chunks);

size_t adjoint_workspace_bytes(int L, int out_len, int K, int chunk, int groups);
int launch_adjoint_fp64(const AdjArgs &a, int max_chunks, hipStream_t stream);
int launch_adjoint_reduce(const AdjReduceArgs &r, hipStream_t stream);
int launch_spline_adjoint(const SplineAdjArgs &a, hipStream_t stream);

}  // namespace few

// Direct (per-mode) interpolated mode summation for gfx950.
//
// Two kernels:
//   * plan_segments_kernel -- the interval -> output-sample map of
//     find_start_inds (src/few/cutils/interpolate.cu:694-723) computed on the
//     device, plus an exclusive prefix of fixed-size sample tiles per
//     interval, so one launch covers every interval (the reference launches
//     once per interval with a device sync after each, interpolate.cu:750-771).
//   * direct_fp64_kernel -- the plain FP64 mode sum of make_waveform
//     (interpolate.cu:444-689): per output sample, for every mode, a cubic
//     amplitude, the phase m*Phi_phi + k*Phi_theta + n*Phi_r, one FP64 sincos
//     and three complex products.  Modes are staged through LDS in chunks,
//     each lane owns SPT samples, the waveform is accumulated into without
//     atomics (every sample belongs to exactly one lane of one block).
//
// Precision policy: the phase pipeline (dense-output evaluation, s, the
// m/k/n combination) is evaluated with contraction off in exactly the
// oracle's operation order, so the argument of sincos is bit-identical to
// the CPU restatement.  Amplitude and complex products may use FMA.
// Mode partial sums are flushed into the output every 5000 modes, matching
// the reference CPU chunking (interpolate.cu:22, 554-600, 682-685).

#include "batch.hpp"
#include "few_common.hpp"
#include "modesum_direct.hpp"

namespace few {

// One workgroup plans one waveform (256 threads).
__device__ void plan_segments_block(const PlanArgs &a)
{
    __shared__ int s_used;
    __shared__ int s_scan[256];
    __shared__ int s_carry;
    const int tid = threadIdx.x;
    const int L = a.L;
    const double T = (a.out_len - 1) * a.dt;
    if (tid == 0) {
        s_used = L;
        s_carry = 0;
    }
    __syncthreads();
    for (int i = 1 + tid; i < L; i += 256)
        if (!(a.h_t[i] < T)) atomicMin(&s_used, i + 1);
    __syncthreads();
    const int used = s_used;
    const int nseg = L - 1;
    for (int base = 0; base < nseg; base += 256) {
        const int j = base + tid;
        int nt = 0;
        if (j < nseg) {
            Segment sg;
            int st, en;
            plan_interval(a.h_t, j, used, L, T, a.dt, a.out_len, &st, &en);
            sg.start = st;
            sg.end = en;
            sg.h_t = a.h_t[j];
            sg.pt = a.pt[j];
            sg.pw = a.pt[j + 1] - a.pt[j];
            sg.hw = a.h_t[j + 1] - a.h_t[j];
            a.segs[j] = sg;
            nt = (en - st + a.tile - 1) / a.tile;
        }
        s_scan[tid] = nt;
        __syncthreads();
        // Hillis-Steele inclusive scan over 256 entries
        for (int off = 1; off < 256; off <<= 1) {
            int v = (tid >= off) ? s_scan[tid - off] : 0;
            __syncthreads();
            s_scan[tid] += v;
            __syncthreads();
        }
        const int carry = s_carry;
        if (j < nseg) a.prefix[j] = carry + s_scan[tid] - nt;
        __syncthreads();
        if (tid == 255) s_carry = carry + s_scan[255];
        __syncthreads();
    }
    if (tid == 0) a.prefix[nseg] = s_carry;
}

__global__ __launch_bounds__(256) void plan_segments_kernel(PlanArgs a) { plan_segments_block(a); }

// Block -> (interval, tile) by binary search over the tile prefix.
__device__ __forceinline__ int find_segment(const int *prefix, int nseg, int b)
{
    int lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
}

#define MCH 256          // modes staged per LDS chunk
#define CPU_CHUNK 5000   // interpolate.cu:22

// One 256*SPT-sample tile (tile q of interval j) of the FP64 direct sum.
template <int SPT>
__device__ void direct_fp64_tile(const DirectArgs &a, int j, int q)
{
    __shared__ double sh[15][MCH];
    const int tid = threadIdx.x;
    const Segment sg = a.segs[j];

    const long np = (long)a.nseg * 3;
    double rp[8], rt[8], rr[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        rp[c] = a.pcoef[c * np + (long)j * 3 + 0];
        rt[c] = a.pcoef[c * np + (long)j * 3 + 1];
        rr[c] = a.pcoef[c * np + (long)j * 3 + 2];
    }

    int idx[SPT];
    bool ok[SPT];
    double x[SPT], x2[SPT], x3[SPT], Pp[SPT], Pt[SPT], Pr[SPT];
    double w_re[SPT], w_im[SPT], tr_re[SPT], tr_im[SPT];
#pragma unroll
    for (int u = 0; u < SPT; u++) {
        idx[u] = sg.start + q * (256 * SPT) + u * 256 + tid;
        ok[u] = idx[u] < sg.end;
        SampleState st;
        sample_state(a.dt, idx[u], sg.h_t, sg.pt, sg.pw, rp, rt, rr, st);
        x[u] = st.x; x2[u] = st.x2; x3[u] = st.x3;
        Pp[u] = st.Pp; Pt[u] = st.Pt; Pr[u] = st.Pr;
        w_re[u] = 0.0; w_im[u] = 0.0;
        if (ok[u]) {
            const double2 w = a.wave[idx[u]];
            w_re[u] = w.x;
            w_im[u] = w.y;
        }
        tr_re[u] = 0.0;
        tr_im[u] = 0.0;
    }

    const int K = a.K;
    const long nb = (long)a.L * 2 * K;
    for (int c0 = 0; c0 < K; c0 += MCH) {
        const int cnt = (K - c0) < MCH ? (K - c0) : MCH;
        __syncthreads();
        if (tid < cnt) {
            const int mode = c0 + tid;
            const long br = (long)j * 2 * K + mode;
            const long bi = br + K;
            sh[0][tid] = a.interp[br];
            sh[1][tid] = a.interp[nb + br];
            sh[2][tid] = a.interp[2 * nb + br];
            sh[3][tid] = a.interp[3 * nb + br];
            sh[4][tid] = a.interp[bi];
            sh[5][tid] = a.interp[nb + bi];
            sh[6][tid] = a.interp[2 * nb + bi];
            sh[7][tid] = a.interp[3 * nb + bi];
            const double2 yp = a.ylm[mode];
            const double2 ym = a.ylm[K + mode];
            sh[8][tid] = yp.x;
            sh[9][tid] = yp.y;
            sh[10][tid] = ym.x;
            sh[11][tid] = ym.y;
            sh[12][tid] = (double)a.m[mode];
            sh[13][tid] = (double)a.k[mode];
            sh[14][tid] = (double)a.n[mode];
        }
        __syncthreads();
        for (int jj = 0; jj < cnt; jj++) {
            const int mode = c0 + jj;
            if (mode > 0 && (mode % CPU_CHUNK) == 0) {
#pragma unroll
                for (int u = 0; u < SPT; u++) {
                    w_re[u] = w_re[u] + tr_re[u];
                    w_im[u] = w_im[u] + tr_im[u];
                    tr_re[u] = 0.0;
                    tr_im[u] = 0.0;
                }
            }
            ModeCoef mc;
            mc.ry = sh[0][jj]; mc.rc1 = sh[1][jj]; mc.rc2 = sh[2][jj]; mc.rc3 = sh[3][jj];
            mc.iy = sh[4][jj]; mc.ic1 = sh[5][jj]; mc.ic2 = sh[6][jj]; mc.ic3 = sh[7][jj];
            mc.yp_re = sh[8][jj]; mc.yp_im = sh[9][jj];
            mc.ym_re = sh[10][jj]; mc.ym_im = sh[11][jj];
            mc.m = sh[12][jj]; mc.k = sh[13][jj]; mc.n = sh[14][jj];
#pragma unroll
            for (int u = 0; u < SPT; u++)
                mode_term_fp64(mc, x[u], x2[u], x3[u], Pp[u], Pt[u], Pr[u], tr_re[u], tr_im[u]);
        }
    }
#pragma unroll
    for (int u = 0; u < SPT; u++) {
        if (ok[u]) {
            double2 w;
            w.x = w_re[u] + tr_re[u];
            w.y = w_im[u] + tr_im[u];
            a.wave[idx[u]] = w;
        }
    }
}

template <int SPT>
__global__ __launch_bounds__(256) void direct_fp64_kernel(DirectArgs a)
{
    const int b = blockIdx.x;
    const int total = a.prefix[a.nseg];
    if (b >= total) return;
    const int j = find_segment(a.prefix, a.nseg, b);
    direct_fp64_tile<SPT>(a, j, b - a.prefix[j]);
}

// ------------------------------------------------------------------ batch
__global__ __launch_bounds__(256) void batch_plan_kernel(BatchDev B, BatchRunDev r, BatchPtrs p,
                                                         int tile)
{
    const WaveDesc d = B.desc[blockIdx.x];
    const double *t = p.t + d.t_off;
    PlanArgs a{t, t, d.L, d.dt, d.N, tile, r.segs + d.seg_off, r.tpref + d.tpref_off};
    plan_segments_block(a);
}

// tiles before each waveform (one workgroup)
__global__ __launch_bounds__(256) void batch_tile_scan_kernel(BatchDev B, BatchRunDev r)
{
    __shared__ int s_scan[256];
    __shared__ int s_carry;
    const int tid = threadIdx.x;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (int base = 0; base < B.B; base += 256) {
        const int b = base + tid;
        int nt = 0;
        if (b < B.B) {
            const WaveDesc d = B.desc[b];
            nt = r.tpref[d.tpref_off + d.L - 1];
        }
        s_scan[tid] = nt;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const int v = (tid >= off) ? s_scan[tid - off] : 0;
            __syncthreads();
            s_scan[tid] += v;
            __syncthreads();
        }
        const int carry = s_carry;
        if (b < B.B) r.wt_prefix[b] = carry + s_scan[tid] - nt;
        __syncthreads();
        if (tid == 255) s_carry = carry + s_scan[255];
        __syncthreads();
    }
    if (tid == 0) r.wt_prefix[B.B] = s_carry;
}

int launch_batch_plan(const BatchDev &B, const BatchRunDev &r, const BatchPtrs &p, int tile,
                      hipStream_t stream)
{
    if (B.B <= 0) return FEW_OK;
    hipLaunchKernelGGL(batch_plan_kernel, dim3(B.B), dim3(256), 0, stream, B, r, p, tile);
    FEW_HIP(hipGetLastError());
    hipLaunchKernelGGL(batch_tile_scan_kernel, dim3(1), dim3(256), 0, stream, B, r);
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

// FP64 control run over the batch: block -> (waveform, interval, tile).
template <int SPT>
__global__ __launch_bounds__(256) void batch_direct_fp64_kernel(BatchDev B, BatchRunDev r,
                                                                BatchPtrs p)
{
    const int g = blockIdx.x;
    if (g >= r.wt_prefix[B.B]) return;
    const int b = upper_index(r.wt_prefix, B.B, g);
    const WaveDesc d = B.desc[b];
    const int *pref = r.tpref + d.tpref_off;
    const int lt = g - r.wt_prefix[b];
    const int j = upper_index(pref, d.L - 1, lt);
    DirectArgs a;
    a.wave = p.wave + d.wave_off / 2;
    a.interp = p.interp + d.interp_off;
    a.pcoef = p.pcoef + d.pc_off;
    a.m = B.m + d.mode_off;
    a.k = B.k + d.mode_off;
    a.n = B.n + d.mode_off;
    a.ylm = reinterpret_cast<const double2 *>(p.ylm + d.ylm_off);
    a.segs = r.segs + d.seg_off;
    a.prefix = pref;
    a.nseg = d.L - 1;
    a.L = d.L;
    a.K = d.K;
    a.dt = d.dt;
    direct_fp64_tile<SPT>(a, j, lt - pref[j]);
}

int launch_batch_direct_fp64(const BatchDev &B, const BatchRunDev &r, const BatchPtrs &p,
                             int max_blocks, hipStream_t stream)
{
    if (max_blocks <= 0) return FEW_OK;
    hipLaunchKernelGGL(batch_direct_fp64_kernel<DIRECT_SPT>, dim3(max_blocks), dim3(256), 0, stream,
                       B, r, p);
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

int launch_plan(const PlanArgs &p, hipStream_t stream)
{
    if (p.L < 2) {
        set_error("plan: need at least 2 knots (got %d)", p.L);
        return FEW_ERR_ARG;
    }
    hipLaunchKernelGGL(plan_segments_kernel, dim3(1), dim3(256), 0, stream, p);
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

int direct_tile(int /*precision*/) { return 256 * DIRECT_SPT; }

int launch_direct_fp64(const DirectArgs &a, int max_blocks, hipStream_t stream)
{
    if (max_blocks <= 0) return FEW_OK;
    hipLaunchKernelGGL(direct_fp64_kernel<DIRECT_SPT>, dim3(max_blocks), dim3(256), 0, stream, a);
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

}  // namespace few

// Host side of the batched mode sum: batch plan (layout of the packed
// buffers, union mode grid, per-waveform gather lists) and the C ABI that
// runs the staged batch pipeline.  No reference counterpart: the reference
// sums one waveform per call (src/few/cutils/interpolate.cu:726-797) and
// batches only sequentially in Python (src/few/waveform/base.py:256-268).

#include <string.h>

#include <map>
#include <vector>

#include "batch.hpp"
#include "few_rocm.h"
#include "modesum_direct.hpp"

namespace few {

struct BatchPlan {
    int B = 0;
    std::vector<WaveDesc> desc;
    FactPlanHost uni;
    long tot[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // t, teuk, pk, ylm, interp, pcoef, wave, scratch (doubles)
    int total_intervals = 0;
    long total_systems = 0;
    int total_tpref = 0;
    long direct_blocks = 0;      // upper bound of FP64 control tiles
    long fact16_blocks = 0;      // upper bound of mixed16 tiles (8-wave tiles)
    void *dev_buf = nullptr;
    BatchDev dev{};
    int device = -1;
};

static inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// Workspace layout of a batch run.
struct BatchWs {
    size_t segs, tpref, wt, region, total;
};
static BatchWs batch_ws_layout(const BatchPlan &p)
{
    BatchWs w;
    size_t off = 0;
    w.segs = off; off = al256(off + sizeof(Segment) * (size_t)(p.total_intervals > 0 ? p.total_intervals : 1));
    w.tpref = off; off = al256(off + sizeof(int) * (size_t)(p.total_tpref > 0 ? p.total_tpref : 1));
    w.wt = off; off = al256(off + sizeof(int) * (size_t)(p.B + 1));
    w.region = off;
    const size_t scratch = sizeof(double) * (size_t)p.tot[7];
    const size_t tiles = p.uni.ok ? fact16_tile_bytes(p.uni) * (size_t)p.total_intervals : 0;
    off = al256(off + (scratch > tiles ? scratch : tiles));
    w.total = off;
    return w;
}

static bool batch_fact16_ok(const BatchPlan &p)
{
    return p.uni.ok && fact16_tile_bytes(p.uni) + 24 * sizeof(double) <= 160 * 1024;
}

}  // namespace few

using namespace few;

extern "C" {

void *few_batch_create(int B, const int *h_L, const int *h_K, const int *h_N, const double *h_dt,
                       const int *h_m, const int *h_k, const int *h_n)
{
    if (B <= 0 || !h_L || !h_K || !h_N || !h_dt || !h_m || !h_k || !h_n) {
        set_error("batch_create: need B > 0 and host arrays");
        return nullptr;
    }
    BatchPlan *bp = new BatchPlan();
    bp->B = B;
    long Ksum = 0;
    for (int b = 0; b < B; b++) {
        if (h_L[b] < 3 || h_K[b] <= 0 || h_N[b] < 0 || !(h_dt[b] > 0.0)) {
            set_error("batch_create: waveform %d has invalid sizes (L=%d K=%d N=%d dt=%g)", b, h_L[b],
                      h_K[b], h_N[b], h_dt[b]);
            delete bp;
            return nullptr;
        }
        Ksum += h_K[b];
    }
    // union geometry over every waveform's modes
    build_fact_plan(h_m, h_k, h_n, (int)Ksum, bp->uni);
    FactPlanHost &u = bp->uni;
    std::map<std::pair<int, int>, int> rowmap;
    for (int r = 0; r < u.Nr; r++) rowmap[{u.row_m[r], u.row_k[r]}] = r;
    const int ncell = u.Nr * u.Nc;
    std::vector<int> cells, ents;
    std::vector<long> sys_prefix(B + 1, 0);
    std::vector<int> seg_prefix(B + 1, 0);
    WaveDesc d{};
    long mode_off = 0;
    for (int b = 0; b < B; b++) {
        const int L = h_L[b], K = h_K[b], N = h_N[b];
        d.L = L;
        d.K = K;
        d.N = N;
        d.dt = h_dt[b];
        d.mode_off = (int)mode_off;
        d.cell_off = (int)cells.size();
        d.ent_off = (int)ents.size();
        d.t_off = bp->tot[0];
        d.teuk_off = bp->tot[1];
        d.pk_off = bp->tot[2];
        d.ylm_off = bp->tot[3];
        d.interp_off = bp->tot[4];
        d.pc_off = bp->tot[5];
        d.wave_off = bp->tot[6];
        d.sc_off = bp->tot[7];
        d.seg_off = bp->total_intervals;
        d.tpref_off = bp->total_tpref;
        d.sys_off = (int)bp->total_systems;
        // gather lists of this waveform on the union grid (local mode indices)
        const int *m = h_m + mode_off, *k = h_k + mode_off, *n = h_n + mode_off;
        std::vector<int> count(ncell + 1, 0);
        auto cell_of = [&](int mm, int kk, int nn) { return rowmap[{mm, kk}] * u.Nc + (nn - u.n_lo); };
        for (int q = 0; q < K; q++) {
            count[cell_of(m[q], k[q], n[q]) + 1]++;
            if (m[q] != 0) count[cell_of(-m[q], -k[q], -n[q]) + 1]++;
        }
        std::vector<int> cs(ncell + 1, 0);
        for (int c = 0; c < ncell; c++) cs[c + 1] = cs[c] + count[c + 1];
        std::vector<int> e(cs[ncell]);
        std::vector<int> fill(cs.begin(), cs.end() - 1);
        for (int q = 0; q < K; q++) {
            e[fill[cell_of(m[q], k[q], n[q])]++] = 2 * q;
            if (m[q] != 0) e[fill[cell_of(-m[q], -k[q], -n[q])]++] = 2 * q + 1;
        }
        cells.insert(cells.end(), cs.begin(), cs.end());
        ents.insert(ents.end(), e.begin(), e.end());
        bp->desc.push_back(d);
        sys_prefix[b] = bp->total_systems;
        seg_prefix[b] = bp->total_intervals;
        bp->tot[0] += L;
        bp->tot[1] += 2L * L * K;
        bp->tot[2] += 3L * L;
        bp->tot[3] += 4L * K;
        bp->tot[4] += 4L * L * 2 * K;
        bp->tot[5] += 8L * (L - 1) * 3;
        bp->tot[6] += 2L * N;
        bp->tot[7] += batch_spline_scratch(L, K);
        bp->total_intervals += L - 1;
        bp->total_tpref += L;
        bp->total_systems += 2L * K + 3;
        bp->direct_blocks += ((long)N + 511) / 512 + (L - 1);
        bp->fact16_blocks += ((long)N + 255) / 256 + (L - 1);
        mode_off += K;
    }
    sys_prefix[B] = bp->total_systems;
    seg_prefix[B] = bp->total_intervals;
    if (cells.empty()) cells.push_back(0);
    if (ents.empty()) ents.push_back(0);

    // one device allocation: desc | sys_prefix | seg_prefix | row_mk | cells | ents | m | k | n
    size_t off = 0;
    const size_t o_desc = off; off = al256(off + sizeof(WaveDesc) * B);
    const size_t o_sys = off; off = al256(off + sizeof(long) * (B + 1));
    const size_t o_seg = off; off = al256(off + sizeof(int) * (B + 1));
    const size_t o_row = off; off = al256(off + sizeof(int) * 2 * (u.Nr > 0 ? u.Nr : 1));
    const size_t o_cell = off; off = al256(off + sizeof(int) * cells.size());
    const size_t o_ent = off; off = al256(off + sizeof(int) * ents.size());
    const size_t o_m = off; off = al256(off + sizeof(int) * Ksum);
    const size_t o_k = off; off = al256(off + sizeof(int) * Ksum);
    const size_t o_n = off; off = al256(off + sizeof(int) * Ksum);
    std::vector<unsigned char> host(off, 0);
    memcpy(&host[o_desc], bp->desc.data(), sizeof(WaveDesc) * B);
    memcpy(&host[o_sys], sys_prefix.data(), sizeof(long) * (B + 1));
    memcpy(&host[o_seg], seg_prefix.data(), sizeof(int) * (B + 1));
    std::vector<int> rowmk;
    for (int r = 0; r < u.Nr; r++) {
        rowmk.push_back(u.row_m[r]);
        rowmk.push_back(u.row_k[r]);
    }
    if (!rowmk.empty()) memcpy(&host[o_row], rowmk.data(), sizeof(int) * rowmk.size());
    memcpy(&host[o_cell], cells.data(), sizeof(int) * cells.size());
    memcpy(&host[o_ent], ents.data(), sizeof(int) * ents.size());
    memcpy(&host[o_m], h_m, sizeof(int) * Ksum);
    memcpy(&host[o_k], h_k, sizeof(int) * Ksum);
    memcpy(&host[o_n], h_n, sizeof(int) * Ksum);
    if (hipGetDevice(&bp->device) != hipSuccess || hipMalloc(&bp->dev_buf, off) != hipSuccess ||
        hipMemcpy(bp->dev_buf, host.data(), off, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("batch_create: device upload of %zu bytes failed", off);
        if (bp->dev_buf) (void)hipFree(bp->dev_buf);
        delete bp;
        return nullptr;
    }
    unsigned char *base = static_cast<unsigned char *>(bp->dev_buf);
    BatchDev &D = bp->dev;
    D.B = B;
    D.total_intervals = bp->total_intervals;
    D.total_systems = bp->total_systems;
    D.desc = reinterpret_cast<const WaveDesc *>(base + o_desc);
    D.sys_prefix = reinterpret_cast<const long *>(base + o_sys);
    D.seg_prefix = reinterpret_cast<const int *>(base + o_seg);
    D.cell_start = reinterpret_cast<const int *>(base + o_cell);
    D.entries = reinterpret_cast<const int *>(base + o_ent);
    D.m = reinterpret_cast<const int *>(base + o_m);
    D.k = reinterpret_cast<const int *>(base + o_k);
    D.n = reinterpret_cast<const int *>(base + o_n);
    FactPlanDev &U = D.uni;
    U.K = 0;
    U.Nr = u.Nr;
    U.Nc = u.Nc;
    U.n_lo = u.n_lo;
    U.RB = u.RB;
    U.KS = u.KS;
    U.CST = u.CST;
    U.rows_arith = u.rows_arith;
    U.m0 = u.m0;
    U.k0 = u.k0;
    U.has_k = u.has_k;
    U.fold = u.fold;
    U.tcols = u.tcols;
    U.tn0 = u.tn0;
    U.row_mk = reinterpret_cast<const int2 *>(base + o_row);
    U.cell_start = nullptr;
    U.entries = nullptr;
    U.mode_cell = nullptr;
    return bp;
}

void few_batch_destroy(void *batch)
{
    if (!batch) return;
    BatchPlan *bp = static_cast<BatchPlan *>(batch);
    if (bp->dev_buf) (void)hipFree(bp->dev_buf);
    delete bp;
}

/* totals10: doubles of t, teuk, phase knots, ylm, interp, pcoef, wave; then
 * spline systems, intervals, supported mask (1: FP64 control, 2: MIXED16).
 * offsets (B x 7, may be NULL): per-waveform double offsets in that order. */
int few_batch_layout(const void *batch, long *totals10, long *offsets)
{
    if (!batch || !totals10) {
        set_error("batch_layout: null argument");
        return FEW_ERR_ARG;
    }
    const BatchPlan *bp = static_cast<const BatchPlan *>(batch);
    for (int i = 0; i < 7; i++) totals10[i] = bp->tot[i];
    totals10[7] = bp->total_systems;
    totals10[8] = bp->total_intervals;
    totals10[9] = 1 | (batch_fact16_ok(*bp) ? 2 : 0);
    if (offsets) {
        for (int b = 0; b < bp->B; b++) {
            const WaveDesc &d = bp->desc[b];
            long *o = offsets + 7L * b;
            o[0] = d.t_off; o[1] = d.teuk_off; o[2] = d.pk_off; o[3] = d.ylm_off;
            o[4] = d.interp_off; o[5] = d.pc_off; o[6] = d.wave_off;
        }
    }
    return FEW_OK;
}

size_t few_batch_workspace_bytes(const void *batch, int precision)
{
    (void)precision;
    if (!batch) return 0;
    return batch_ws_layout(*static_cast<const BatchPlan *>(batch)).total;
}

int few_batch_run_async(const void *batch, const double *t, const double *teuk,
                        const double *phase_knots, const double *ylm, double *interp, double *pcoef,
                        void *wave, int precision, int variant, int stage, void *workspace,
                        size_t workspace_bytes, void *stream)
{
    if (!batch || !t || !teuk || !ylm || !interp || !pcoef || !wave || !workspace) {
        set_error("batch_run: null argument");
        return FEW_ERR_ARG;
    }
    if (stage < 1 || stage > 15 || ((stage & 8) && !(stage & 4))) {
        set_error("batch_run: stage is a mask of 1 (splines), 2 (plan + tiles), 4 (sum), "
                  "8 (with 4: overwrite)");
        return FEW_ERR_ARG;
    }
    const BatchPlan &bp = *static_cast<const BatchPlan *>(batch);
    const BatchWs L = batch_ws_layout(bp);
    if (workspace_bytes < L.total) {
        set_error("batch_run: workspace of %zu bytes, need %zu", workspace_bytes, L.total);
        return FEW_ERR_ARG;
    }
    if (precision != FEW_PRECISION_FP64 && precision != FEW_PRECISION_MIXED16) {
        set_error("batch_run: precision %d not available for batches (use FP64 or MIXED16)", precision);
        return FEW_ERR_UNSUPPORTED;
    }
    if (precision == FEW_PRECISION_MIXED16 && !batch_fact16_ok(bp)) {
        set_error("batch_run: the batch's union mode grid does not fit the mixed16 tiles");
        return FEW_ERR_UNSUPPORTED;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    char *w = static_cast<char *>(workspace);
    BatchRunDev r{reinterpret_cast<Segment *>(w + L.segs), reinterpret_cast<int *>(w + L.tpref),
                  reinterpret_cast<int *>(w + L.wt), 0};
    BatchPtrs p{t, teuk, phase_knots, ylm, interp, pcoef, reinterpret_cast<double *>(w + L.region),
                static_cast<double2 *>(wave)};
    const int tile = precision == FEW_PRECISION_FP64 ? direct_tile(FEW_PRECISION_FP64)
                                                     : fact16_tile(variant);
    if (precision == FEW_PRECISION_MIXED16 && (stage & 1) && !(stage & 2)) {
        // the spline scratch shares the region of the coefficient tiles: what
        // an earlier stage 2 prepared is gone, so a sum needs stage 2 again
        if (stage & 4) {
            set_error("batch_run: stage 1 overwrites the mixed16 coefficient tiles; run stage 2 "
                      "before the sum");
            return FEW_ERR_ARG;
        }
        const PrepKey spent{workspace, batch, -1, 0, 0, 0};
        (void)prep_stage(true, false, spent, "batch_run");
    }
    const PrepKey key{workspace, batch, precision, tile, 0, 0};
    int rc = prep_stage(stage & 2, stage & 4, key, "batch_run");
    if (rc) return rc;
    if (stage & 1) {
        rc = launch_batch_splines(bp.dev, p, s);
        if (rc) return rc;
    }
    if (precision == FEW_PRECISION_FP64) {
        if (stage & 2) {
            rc = launch_batch_plan(bp.dev, r, p, tile, s);
            if (rc) return rc;
        }
        if (stage & 4) {
            if (bp.direct_blocks > 0x7fffffffL) {
                set_error("batch_run: too many tiles");
                return FEW_ERR_LIMIT;
            }
            if (stage & 8) {
                rc = launch_zero_complex(static_cast<double2 *>(wave), bp.tot[6] / 2, s);
                if (rc) return rc;
            }
            return launch_batch_direct_fp64(bp.dev, r, p, (int)bp.direct_blocks, s);
        }
        return FEW_OK;
    }
    if (stage & 2) {
        rc = launch_batch_plan(bp.dev, r, p, tile, s);
        if (rc) return rc;
    }
    unsigned char *tiles = reinterpret_cast<unsigned char *>(w + L.region);
    const long tb = (long)fact16_tile_bytes(bp.uni);
    const int sub = ((stage & 2) ? 1 : 0) | ((stage & 4) ? 2 : 0) | ((stage & 8) ? 4 : 0);
    if (!sub) return FEW_OK;
    return launch_batch_fact16(bp.dev, r, p, tiles, tb, fact16_kb(bp.uni),
                               variant, sub, s);
}

}  // extern "C"

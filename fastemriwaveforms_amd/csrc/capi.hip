// C ABI of the MI355X interpolated mode-summation library.
//
// The entry points mirror the reference's native interface
// (src/few/cutils/interpolate.hh:27-38, bound at src/few/cutils/pyinterp.pyx:10-16)
// with plain pointers and sizes, an explicit HIP stream, and an int status
// instead of exit() on error (the reference aborts the process through
// gpuErrchk / CUSPARSE_CALL, src/few/cutils/global.h:49-57,
// src/few/cutils/interpolate.hh:12-23).  The message of the last failing
// call on the calling thread is available from few_last_error().

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "aak.hpp"
#include "adjoint.hpp"
#include "amp2d.hpp"
#include "eval.hpp"
#include "fd.hpp"
#include "few_common.hpp"
#include "few_rocm.h"
#include "modesum_direct.hpp"
#include "modesum_fact.hpp"

namespace few {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int check_hip(hipError_t e, const char *what)
{
    if (e == hipSuccess) return FEW_OK;
    set_error("%s failed: %s", what, hipGetErrorString(e));
    return FEW_ERR_HIP;
}

int ensure_dynamic_lds(const void *kernel, int bytes)
{
    struct Entry {
        const void *k;
        int dev, bytes;
    };
    static std::mutex mu;
    static Entry table[1024];
    static int n = 0;
    int dev = 0;
    FEW_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    for (int i = 0; i < n; i++) {
        if (table[i].k == kernel && table[i].dev == dev) {
            if (table[i].bytes >= bytes) return FEW_OK;
            FEW_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
            table[i].bytes = bytes;
            return FEW_OK;
        }
    }
    FEW_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    if (n < 1024) table[n++] = Entry{kernel, dev, bytes};
    return FEW_OK;
}

// Library-owned device workspace for the synchronous drop-in entry points
// (one per device, grown on demand, kept until few_release_workspace()).
struct DeviceWorkspace {
    void *ptr = nullptr;
    size_t bytes = 0;
};
static std::mutex g_ws_mu;
static DeviceWorkspace g_ws[64];

static int library_workspace(int dev, size_t bytes, void **out)
{
    if (dev < 0 || dev >= 64) {
        set_error("device index %d out of range", dev);
        return FEW_ERR_ARG;
    }
    DeviceWorkspace &w = g_ws[dev];
    if (w.bytes < bytes) {
        if (w.ptr) {
            FEW_HIP(hipDeviceSynchronize());
            FEW_HIP(hipFree(w.ptr));
            w.ptr = nullptr;
            w.bytes = 0;
        }
        size_t nb = bytes < (1u << 20) ? (1u << 20) : bytes;
        FEW_HIP(hipMalloc(&w.ptr, nb));
        w.bytes = nb;
    }
    *out = w.ptr;
    return FEW_OK;
}

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Layout of the mode-sum planning workspace.
struct PlanLayout {
    size_t h_t, pt, segs, prefix, total;
};
static PlanLayout plan_layout(int L)
{
    PlanLayout p;
    size_t off = 0;
    p.h_t = off; off = align256(off + sizeof(double) * (size_t)L);
    p.pt = off; off = align256(off + sizeof(double) * (size_t)L);
    p.segs = off; off = align256(off + sizeof(Segment) * (size_t)(L > 1 ? L - 1 : 1));
    p.prefix = off; off = align256(off + sizeof(int) * (size_t)L);
    p.total = off;
    return p;
}

// Workspace of the factorised path: plan layout, then the coefficient tiles.
static size_t fact_coef_offset(int L) { return plan_layout(L).total; }
static size_t fact_tiles_bytes(const FactPlanHost &p, int L, int precision)
{
    const size_t n = (size_t)(L > 1 ? L - 1 : 1);
    if (precision == FEW_PRECISION_MIXED16) return align256(fact16_tile_bytes(p) * n);
    const size_t elem = precision == FEW_PRECISION_FACT64 ? sizeof(double) : sizeof(float);
    return align256(elem * fact_coef_floats_per_interval(p) * n);
}
// plan layout | coefficient tiles of the precision
static size_t fact_ws_bytes(const FactPlanHost &p, int L, int precision)
{
    return fact_coef_offset(L) + fact_tiles_bytes(p, L, precision);
}
static bool fact16_supported(const FactPlanHost &p)
{
    return p.ok && fact16_tile_bytes(p) + 24 * sizeof(double) <= 160 * 1024;
}
static bool fact64_supported(const FactPlanHost &p)
{
    return p.ok && fact_coef_floats_per_interval(p) * sizeof(double) + 24 * sizeof(double) <=
                       160 * 1024;
}

struct DeviceGuard {
    int prev = -1;
    int ok = FEW_OK;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (dev >= 0 && dev != prev) ok = check_hip(hipSetDevice(dev), "hipSetDevice");
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

static int check_mode_args(const void *wave, const double *interp, const double *pcoef, const int *m,
                           const int *k, const int *n, int init_len, int out_len, int K,
                           const void *ylm, double dt)
{
    if (!wave || !interp || !pcoef || !m || !k || !n || !ylm) {
        set_error("get_waveform: null pointer argument");
        return FEW_ERR_ARG;
    }
    if (init_len < 2 || out_len < 0 || K < 0 || !(dt > 0.0)) {
        set_error("get_waveform: invalid sizes (init_len=%d out_len=%d modes=%d dt=%g)", init_len,
                  out_len, K, dt);
        return FEW_ERR_ARG;
    }
    return FEW_OK;
}

// Zero fill of a complex128 range (a kernel rather than hipMemsetAsync, so
// the call stays capturable in a graph).
__global__ __launch_bounds__(256) void zero_complex_kernel(double2 *w, long n)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        w[i] = make_double2(0.0, 0.0);
}

int launch_zero_complex(double2 *w, long n, hipStream_t stream)
{
    if (n <= 0) return FEW_OK;
    const long blocks = std::min<long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(zero_complex_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, w, n);
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

// What the last preparing call left in each workspace (see few_common.hpp).
// Keyed by workspace address; bounded, cleared when full.
static std::mutex g_prep_mu;
static std::vector<PrepKey> g_prep;

static void prep_record(const PrepKey &k)
{
    std::lock_guard<std::mutex> lk(g_prep_mu);
    for (PrepKey &e : g_prep)
        if (e.ws == k.ws) {
            e = k;
            return;
        }
    if (g_prep.size() >= 4096) g_prep.clear();
    g_prep.push_back(k);
}

static int prep_check(const PrepKey &k, const char *what)
{
    std::lock_guard<std::mutex> lk(g_prep_mu);
    for (const PrepKey &e : g_prep)
        if (e.ws == k.ws) {
            if (e.plan == k.plan && e.precision == k.precision && e.tile == k.tile &&
                e.L == k.L && e.out_len == k.out_len)
                return FEW_OK;
            set_error("%s: sum with precision %d, %d samples per tile, L %d, %d samples%s; the "
                      "workspace was prepared with precision %d, %d samples per tile, L %d, "
                      "%d samples",
                      what, k.precision, k.tile, k.L, k.out_len,
                      e.plan == k.plan ? "" : " and another plan", e.precision, e.tile, e.L,
                      e.out_len);
            return FEW_ERR_ARG;
        }
    set_error("%s: sum on a workspace no preparing call prepared", what);
    return FEW_ERR_ARG;
}

int prep_stage(bool prepares, bool sums, const PrepKey &k, const char *what)
{
    if (prepares) {
        prep_record(k);
        return FEW_OK;
    }
    return sums ? prep_check(k, what) : FEW_OK;
}

// Plans (device) and runs one waveform's mode sum on `stream`; h_t and pt
// are device pointers.  stage: 1 = plan (+ coefficient tiles), 2 = sum,
// 4 (with 2) = overwrite: the output is written rather than added to, and
// samples no interval covers are zeroed.
static int run_waveform(double2 *wave, const double *interp, const double *pt, const double *pcoef,
                        const int *m, const int *k, const int *n, int L, int out_len, int K,
                        const double2 *ylm, double dt, const double *h_t, int precision, char *ws,
                        size_t ws_bytes, hipStream_t stream, const FactPlan *plan = nullptr,
                        int variant = 0, int stage = 3)
{
    if (out_len == 0) return FEW_OK;
    if (K == 0) {
        // no modes: h = 0, which the overwriting sum must still write
        return ((stage & 2) && (stage & 4)) ? launch_zero_complex(wave, out_len, stream) : FEW_OK;
    }
    const PlanLayout pl = plan_layout(L);
    if (ws_bytes < pl.total) {
        set_error("get_waveform: workspace of %zu bytes, need %zu", ws_bytes, pl.total);
        return FEW_ERR_ARG;
    }
    Segment *segs = reinterpret_cast<Segment *>(ws + pl.segs);
    int *prefix = reinterpret_cast<int *>(ws + pl.prefix);
    if (precision == FEW_PRECISION_MIXED16) {
        if (!plan || !fact16_supported(plan->host)) {
            set_error("get_waveform: mixed16 precision needs a supported mode plan");
            return FEW_ERR_UNSUPPORTED;
        }
        const FactPlanHost &ph = plan->host;
        if (ph.K != K) {
            set_error("get_waveform: plan built for %d modes, called with %d", ph.K, K);
            return FEW_ERR_ARG;
        }
        if (ws_bytes < fact_ws_bytes(ph, L, precision)) {
            set_error("get_waveform: workspace of %zu bytes, need %zu", ws_bytes,
                      fact_ws_bytes(ph, L, precision));
            return FEW_ERR_ARG;
        }
        unsigned char *tiles = reinterpret_cast<unsigned char *>(ws + fact_coef_offset(L));
        const long tb = (long)fact16_tile_bytes(ph);
        const int tile = fact16_tile(variant);
        const long max_blocks = ((long)out_len + tile - 1) / tile + (L - 1);
        if (max_blocks > 0x7fffffffL) {
            set_error("get_waveform: output too long (%d samples)", out_len);
            return FEW_ERR_LIMIT;
        }
        const PrepKey key{ws, plan, precision, tile, L, out_len};
        if (int rc = prep_stage(stage & 1, stage & 2, key, "get_waveform")) return rc;
        if (stage & 1) {
            PlanArgs p{h_t, pt, L, dt, out_len, tile, segs, prefix};
            int rc = launch_plan(p, stream);
            if (rc) return rc;
            Fact16PrepArgs pa{interp, ylm, h_t, L, tiles, tb, fact16_kb(ph), ph.RB,
                              plan->dev};
            rc = launch_fact16_prep(pa, stream);
            if (rc) return rc;
        }
        if (!(stage & 2)) return FEW_OK;
        Fact16Args fa{wave, pcoef, segs, prefix, L - 1, dt, tiles, tb, fact16_kb(ph),
                      plan->dev};
        fa.overwrite = (stage & 4) ? 1 : 0;
        fa.out_len = out_len;
        return launch_fact16_sum(fa, ph.RB, (int)max_blocks, variant, stream);
    }
    if (precision == FEW_PRECISION_MIXED) {
        if (!plan || !plan->host.ok) {
            set_error("get_waveform: mixed precision needs a supported mode plan");
            return FEW_ERR_UNSUPPORTED;
        }
        const FactPlanHost &ph = plan->host;
        if (ph.K != K) {
            set_error("get_waveform: plan built for %d modes, called with %d", ph.K, K);
            return FEW_ERR_ARG;
        }
        if (ws_bytes < fact_ws_bytes(ph, L, precision)) {
            set_error("get_waveform: workspace of %zu bytes, need %zu", ws_bytes,
                      fact_ws_bytes(ph, L, precision));
            return FEW_ERR_ARG;
        }
        float *C = reinterpret_cast<float *>(ws + fact_coef_offset(L));
        const long c_stride = (long)fact_coef_floats_per_interval(ph);
        const int tile = fact_tile(ph, variant);
        const long max_blocks = ((long)out_len + tile - 1) / tile + (L - 1);
        if (max_blocks > 0x7fffffffL) {
            set_error("get_waveform: output too long (%d samples)", out_len);
            return FEW_ERR_LIMIT;
        }
        const PrepKey key{ws, plan, precision, tile, L, out_len};
        if (int rc = prep_stage(stage & 1, stage & 2, key, "get_waveform")) return rc;
        if (stage & 1) {
            PlanArgs p{h_t, pt, L, dt, out_len, tile, segs, prefix};
            int rc = launch_plan(p, stream);
            if (rc) return rc;
            FactPrepArgs pa{interp, ylm, h_t, L, C, c_stride, plan->dev};
            rc = launch_fact_prep(pa, stream);
            if (rc) return rc;
        }
        if (!(stage & 2)) return FEW_OK;
        if (stage & 4) {
            int rc = launch_zero_complex(wave, out_len, stream);
            if (rc) return rc;
        }
        FactArgs fa{wave, pcoef, segs, prefix, L - 1, dt, C, c_stride, plan->dev};
        return launch_fact_sum(fa, ph, (int)max_blocks, variant, stream);
    }
    if (precision == FEW_PRECISION_FACT64) {
        if (!plan || !fact64_supported(plan->host)) {
            set_error("get_waveform: fact64 precision needs a supported mode plan");
            return FEW_ERR_UNSUPPORTED;
        }
        const FactPlanHost &ph = plan->host;
        if (ph.K != K) {
            set_error("get_waveform: plan built for %d modes, called with %d", ph.K, K);
            return FEW_ERR_ARG;
        }
        if (ws_bytes < fact_ws_bytes(ph, L, precision)) {
            set_error("get_waveform: workspace of %zu bytes, need %zu", ws_bytes,
                      fact_ws_bytes(ph, L, precision));
            return FEW_ERR_ARG;
        }
        double *C = reinterpret_cast<double *>(ws + fact_coef_offset(L));
        const long c_stride = (long)fact_coef_floats_per_interval(ph);
        if (variant != 0 && variant != 4) {
            set_error("get_waveform: fact64 variant %d unknown (0 or 4)", variant);
            return FEW_ERR_ARG;
        }
        const int tile = fact64_tile(variant);
        const long max_blocks = ((long)out_len + tile - 1) / tile + (L - 1);
        if (max_blocks > 0x7fffffffL) {
            set_error("get_waveform: output too long (%d samples)", out_len);
            return FEW_ERR_LIMIT;
        }
        const PrepKey key{ws, plan, precision, tile, L, out_len};
        if (int rc = prep_stage(stage & 1, stage & 2, key, "get_waveform")) return rc;
        if (stage & 1) {
            PlanArgs p{h_t, pt, L, dt, out_len, tile, segs, prefix};
            int rc = launch_plan(p, stream);
            if (rc) return rc;
            FactPrepArgs pa{interp, ylm, h_t, L, nullptr, c_stride, plan->dev};
            pa.C64 = C;
            rc = launch_fact_prep(pa, stream);
            if (rc) return rc;
        }
        if (!(stage & 2)) return FEW_OK;
        FactArgs fa{wave, pcoef, segs, prefix, L - 1, dt, nullptr, c_stride, plan->dev};
        fa.C64 = C;
        fa.overwrite = (stage & 4) ? 1 : 0;
        fa.out_len = out_len;
        return launch_fact64_sum(fa, ph, (int)max_blocks, variant, stream);
    }
    if (precision != FEW_PRECISION_FP64) {
        set_error("get_waveform: precision %d is not supported", precision);
        return FEW_ERR_UNSUPPORTED;
    }
    const int tile = direct_tile(precision);
    const PrepKey key{ws, plan, precision, tile, L, out_len};
    if (int rc = prep_stage(stage & 1, stage & 2, key, "get_waveform")) return rc;
    if (stage & 1) {
        PlanArgs p{h_t, pt, L, dt, out_len, tile, segs, prefix};
        int rc = launch_plan(p, stream);
        if (rc) return rc;
    }
    if (!(stage & 2)) return FEW_OK;
    if (stage & 4) {
        int rc = launch_zero_complex(wave, out_len, stream);
        if (rc) return rc;
    }
    DirectArgs a;
    a.wave = wave;
    a.interp = interp;
    a.pcoef = pcoef;
    a.m = m;
    a.k = k;
    a.n = n;
    a.ylm = ylm;
    a.segs = segs;
    a.prefix = prefix;
    a.nseg = L - 1;
    a.L = L;
    a.K = K;
    a.dt = dt;
    const long max_blocks = ((long)out_len + tile - 1) / tile + (L - 1);
    if (max_blocks > 0x7fffffffL) {
        set_error("get_waveform: output too long (%d samples)", out_len);
        return FEW_ERR_LIMIT;
    }
    return launch_direct_fp64(a, (int)max_blocks, stream);
}

// ------------------------------------------------------------------ adjoint

#define ADJ_CHUNK 2048   // samples per chunk of the mode-sum adjoint
#define ADJ_GROUP 128    // modes per workgroup of the mode-sum adjoint

struct AdjLayout {
    size_t segs, prefix, amp, phase, total;
    long max_chunks;
    int groups;
};
static AdjLayout adj_layout(int L, int out_len, int K)
{
    AdjLayout a;
    a.max_chunks = ((long)out_len + ADJ_CHUNK - 1) / ADJ_CHUNK + (L > 1 ? L - 1 : 1);
    a.groups = (K + ADJ_GROUP - 1) / ADJ_GROUP;
    size_t off = 0;
    a.segs = off; off = align256(off + sizeof(Segment) * (size_t)(L > 1 ? L - 1 : 1));
    a.prefix = off; off = align256(off + sizeof(int) * (size_t)L);
    a.amp = off; off = align256(off + sizeof(double) * (size_t)a.max_chunks * 8 * (size_t)K);
    a.phase = off; off = align256(off + sizeof(double) * (size_t)a.max_chunks * a.groups * 12);
    a.total = off;
    return a;
}

// Mixed-precision adjoint: shapes and workspace.
struct Adj16Layout {
    int RBM = 0, NBC = 0, groups = 0;
    int NBCM = 0;                  // column blocks of the interval moments M (ceil(Nc / 16))
    long max_chunks = 0, tile = 0;  // tile: RBM*16 x NBCM*16, one interval's M
    long tile_part = 0;            // RBM*16 x NBC*16, one chunk's moments
    size_t segs = 0, prefix = 0, part = 0, M = 0, phase = 0, total = 0;
    bool ok = false;
};
static Adj16Layout adj16_layout(const FactPlanHost &h, int L, int out_len)
{
    Adj16Layout a;
    a.RBM = (2 * A16_MOM * h.Nr + 15) / 16;
    a.NBC = (h.Nc + 15) / 16;
    a.NBCM = a.NBC;
    a.groups = (h.K + 255) / 256;
    a.tile = (long)a.RBM * 16 * a.NBC * 16;
    a.tile_part = a.tile;
    a.max_chunks = ((long)out_len + A16_CHUNK - 1) / A16_CHUNK + (L > 1 ? L - 1 : 1);
    size_t off = 0;
    a.segs = off; off = align256(off + sizeof(Segment) * (size_t)(L > 1 ? L - 1 : 1));
    a.prefix = off; off = align256(off + sizeof(int) * (size_t)L);
    a.part = off; off = align256(off + sizeof(float) * (size_t)a.max_chunks * a.tile);
    a.M = off; off = align256(off + sizeof(double) * (size_t)(L > 1 ? L - 1 : 1) * a.tile);
    a.phase = off; off = align256(off + sizeof(double) * (size_t)L * a.groups * 12);
    a.total = off;
    a.ok = h.ok && a.RBM >= 1 && a.RBM <= 24 && a.NBC >= 1 && a.NBC <= 8 &&
           adjoint16_lds_bytes(a.RBM, a.NBC, h.Nr) <= 160 * 1024;
    return a;
}

// FP64 factorised adjoint: the split-FP16 adjoint's shapes, FP64 chunk moments;
// with folded tiles the chunk moments have 2 tcols real columns (n >= 0).
static Adj16Layout adj64_layout(const FactPlanHost &h, int L, int out_len)
{
    Adj16Layout a;
    a.RBM = (2 * A16_MOM * h.Nr + 15) / 16;
    a.NBCM = (h.Nc + 15) / 16;
    a.NBC = h.fold ? (2 * h.tcols + 15) / 16 : a.NBCM;
    a.groups = (h.K + 255) / 256;
    a.tile = (long)a.RBM * 16 * a.NBCM * 16;
    a.tile_part = (long)a.RBM * 16 * a.NBC * 16;
    a.max_chunks = ((long)out_len + A64_CHUNK - 1) / A64_CHUNK + (L > 1 ? L - 1 : 1);
    size_t off = 0;
    a.segs = off; off = align256(off + sizeof(Segment) * (size_t)(L > 1 ? L - 1 : 1));
    a.prefix = off; off = align256(off + sizeof(int) * (size_t)L);
    a.part = off; off = align256(off + sizeof(double) * (size_t)a.max_chunks * a.tile_part);
    a.M = off; off = align256(off + sizeof(double) * (size_t)(L > 1 ? L - 1 : 1) * a.tile);
    a.phase = off; off = align256(off + sizeof(double) * (size_t)L * a.groups * 12);
    a.total = off;
    a.ok = h.ok && a.RBM >= 1 && a.RBM <= 24 && a.NBC >= 1 && a.NBC <= 8 &&
           adjoint64_lds_bytes(a.RBM, a.NBC) <= 160 * 1024;
    return a;
}


}  // namespace few

using namespace few;

extern "C" {

const char *few_last_error(void) { return g_err; }

int few_abi_version(void) { return FEW_ROCM_ABI_VERSION; }

int few_interpolate_arrays(double *t_arr, double *interp_array, int ninterps, int length, double *B,
                           double *upper_diag, double *diag, double *lower_diag, void *stream)
{
    if (!t_arr || !interp_array || !B || !upper_diag || !diag || !lower_diag) {
        set_error("interpolate_arrays: null pointer argument");
        return FEW_ERR_ARG;
    }
    if (ninterps <= 0 || length < 3) {
        set_error("interpolate_arrays: need ninterps > 0 and length >= 3 (got %d, %d)", ninterps,
                  length);
        return FEW_ERR_ARG;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    SplineGroup g;
    g.y = interp_array;
    g.y_sys = length;
    g.y_knot = 1;
    g.t = t_arr;
    g.t_sys = length;
    g.t_knot = 1;
    g.out = interp_array;
    g.o_coef = (long)ninterps * length;
    g.o_sys = length;
    g.o_knot = 1;
    g.nsys = ninterps;
    g.dense = 0;
    double *scratch[4] = {diag, upper_diag, lower_diag, B};
    int rc = launch_spline_build(&g, 1, length, scratch, nullptr, s);
    if (rc) return rc;
    FEW_HIP(hipStreamSynchronize(s));
    return FEW_OK;
}

int few_spline_build_async(const double *y, long y_sys, long y_knot, const double *t, long t_sys,
                           long t_knot, double *out, long o_coef, long o_sys, long o_knot, int nsys,
                           int length, int dense, double *scratch, void *stream)
{
    if (!y || !t || !out || !scratch) {
        set_error("spline_build: null pointer argument");
        return FEW_ERR_ARG;
    }
    if (nsys <= 0 || length < 3) {
        set_error("spline_build: need nsys > 0 and length >= 3 (got %d, %d)", nsys, length);
        return FEW_ERR_ARG;
    }
    SplineGroup g{y, y_sys, y_knot, t, t_sys, t_knot, out, o_coef, o_sys, o_knot, nsys, dense ? 1 : 0};
    const long f = (long)length * nsys;
    double *sc[4] = {scratch, scratch + f, scratch + 2 * f, scratch + 3 * f};
    return launch_spline_build(&g, 1, length, sc, nullptr, reinterpret_cast<hipStream_t>(stream));
}

int few_prepare_modes_async(const double *teuk, const double *phase_knots, const double *t,
                            int length, int num_modes, double *interp_out, double *pcoef_out,
                            double *scratch, void *stream)
{
    if (!teuk || !t || !interp_out || !scratch || (phase_knots && !pcoef_out)) {
        set_error("prepare_modes: null pointer argument");
        return FEW_ERR_ARG;
    }
    if (num_modes <= 0 || length < 3) {
        set_error("prepare_modes: need modes > 0 and length >= 3 (got %d, %d)", num_modes, length);
        return FEW_ERR_ARG;
    }
    const long K = num_modes, L = length;
    SplineGroup g[3];
    // Re parts of teuk (L, K) complex -> interp arrays [0, K)
    g[0] = SplineGroup{teuk, 2, 2 * K, t, 0, 1, interp_out, L * 2 * K, 1, 2 * K, num_modes, 0};
    // Im parts -> interp arrays [K, 2K)
    g[1] = SplineGroup{teuk + 1, 2, 2 * K, t, 0, 1, interp_out + K, L * 2 * K, 1, 2 * K, num_modes, 0};
    // phase knots (3, L) -> dense-output coefficients (8, L-1, 3)
    g[2] = SplineGroup{phase_knots, L, 1, t, 0, 1, pcoef_out, 0, 0, 0, 3, 1};
    const int ng = phase_knots ? 3 : 2;
    const long f = L * (2 * K + (phase_knots ? 3 : 0));
    double *sc[4] = {scratch, scratch + f, scratch + 2 * f, scratch + 3 * f};
    return launch_spline_build(g, ng, length, sc, nullptr, reinterpret_cast<hipStream_t>(stream));
}

size_t few_waveform_workspace_bytes(int init_len, int out_len, int num_modes, int precision)
{
    (void)out_len;
    (void)num_modes;
    (void)precision;
    return plan_layout(init_len < 2 ? 2 : init_len).total;
}

int few_get_waveform(void *d_waveform, double *interp_array, double *phase_interp_t,
                     double *phase_interp_coeffs, int *d_m, int *d_k, int *d_n, int init_len,
                     int out_len, int num_teuk_modes, void *d_Ylms, double delta_t, double *h_t,
                     int dev, void *stream)
{
    int rc = check_mode_args(d_waveform, interp_array, phase_interp_coeffs, d_m, d_k, d_n, init_len,
                             out_len, num_teuk_modes, d_Ylms, delta_t);
    if (rc) return rc;
    if (!phase_interp_t || !h_t) {
        set_error("get_waveform: null host time array");
        return FEW_ERR_ARG;
    }
    DeviceGuard guard(dev);
    if (guard.ok) return guard.ok;
    int cur = 0;
    FEW_HIP(hipGetDevice(&cur));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lk(g_ws_mu);
    const PlanLayout pl = plan_layout(init_len);
    void *ws = nullptr;
    rc = library_workspace(cur, pl.total, &ws);
    if (rc) return rc;
    char *w = static_cast<char *>(ws);
    // h_t and phase_interp_t are host arrays in the reference interface
    // (interpolate.cu:739, :767-768); stage them on the device.
    FEW_HIP(hipMemcpyAsync(w + pl.h_t, h_t, sizeof(double) * init_len, hipMemcpyHostToDevice, s));
    FEW_HIP(hipMemcpyAsync(w + pl.pt, phase_interp_t, sizeof(double) * init_len,
                           hipMemcpyHostToDevice, s));
    rc = run_waveform(static_cast<double2 *>(d_waveform), interp_array,
                      reinterpret_cast<double *>(w + pl.pt), phase_interp_coeffs, d_m, d_k, d_n,
                      init_len, out_len, num_teuk_modes, static_cast<const double2 *>(d_Ylms),
                      delta_t, reinterpret_cast<double *>(w + pl.h_t), FEW_PRECISION_FP64, w,
                      pl.total, s);
    if (rc) return rc;
    FEW_HIP(hipStreamSynchronize(s));
    return FEW_OK;
}

int few_get_waveform_async(void *d_waveform, const double *interp_array, const double *d_phase_interp_t,
                           const double *phase_interp_coeffs, const int *d_m, const int *d_k,
                           const int *d_n, int init_len, int out_len, int num_teuk_modes,
                           const void *d_Ylms, double delta_t, const double *d_h_t, int precision,
                           void *workspace, size_t workspace_bytes, void *stream)
{
    int rc = check_mode_args(d_waveform, interp_array, phase_interp_coeffs, d_m, d_k, d_n, init_len,
                             out_len, num_teuk_modes, d_Ylms, delta_t);
    if (rc) return rc;
    if (!d_phase_interp_t || !d_h_t || !workspace) {
        set_error("get_waveform_async: null device time array or workspace");
        return FEW_ERR_ARG;
    }
    char *w = static_cast<char *>(workspace);
    return run_waveform(static_cast<double2 *>(d_waveform), interp_array, d_phase_interp_t,
                        phase_interp_coeffs, d_m, d_k, d_n, init_len, out_len, num_teuk_modes,
                        static_cast<const double2 *>(d_Ylms), delta_t, d_h_t, precision, w,
                        workspace_bytes, reinterpret_cast<hipStream_t>(stream));
}

void *few_mode_plan_create(const int *h_m, const int *h_k, const int *h_n, int num_modes)
{
    if (!h_m || !h_k || !h_n || num_modes <= 0) {
        set_error("mode_plan_create: need host m/k/n arrays and modes > 0");
        return nullptr;
    }
    FactPlan *pl = new FactPlan();
    build_fact_plan(h_m, h_k, h_n, num_modes, pl->host);
    FactPlanHost &h = pl->host;
    pl->dev.K = h.K;
    pl->dev.Nr = h.Nr;
    pl->dev.Nc = h.Nc;
    pl->dev.n_lo = h.n_lo;
    pl->dev.RB = h.RB;
    pl->dev.KS = h.KS;
    pl->dev.CST = h.CST;
    pl->dev.rows_arith = h.rows_arith;
    pl->dev.m0 = h.m0;
    pl->dev.k0 = h.k0;
    pl->dev.has_k = h.has_k;
    pl->dev.fold = h.fold;
    pl->dev.tcols = h.tcols;
    pl->dev.tn0 = h.tn0;
    if (!h.ok) return pl;  // usable for the direct paths; mixed reports unsupported
    std::vector<int> buf;
    for (int r = 0; r < h.Nr; r++) {
        buf.push_back(h.row_m[r]);
        buf.push_back(h.row_k[r]);
    }
    const size_t o_cell = buf.size();
    buf.insert(buf.end(), h.cell_start.begin(), h.cell_start.end());
    const size_t o_ent = buf.size();
    buf.insert(buf.end(), h.entries.begin(), h.entries.end());
    const size_t o_mc = buf.size();
    buf.insert(buf.end(), h.mode_cell.begin(), h.mode_cell.end());
    if (buf.empty()) buf.push_back(0);
    pl->dev_bytes = buf.size() * sizeof(int);
    if (hipGetDevice(&pl->device) != hipSuccess ||
        hipMalloc(&pl->dev_buf, pl->dev_bytes) != hipSuccess ||
        hipMemcpy(pl->dev_buf, buf.data(), pl->dev_bytes, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("mode_plan_create: device upload failed");
        if (pl->dev_buf) (void)hipFree(pl->dev_buf);
        delete pl;
        return nullptr;
    }
    int *d = static_cast<int *>(pl->dev_buf);
    pl->dev.row_mk = reinterpret_cast<const int2 *>(d);
    pl->dev.cell_start = d + o_cell;
    pl->dev.entries = d + o_ent;
    pl->dev.mode_cell = d + o_mc;
    return pl;
}

int few_mode_plan_info(const void *plan, int *info7)
{
    if (!plan || !info7) {
        set_error("mode_plan_info: null argument");
        return FEW_ERR_ARG;
    }
    const FactPlanHost &h = static_cast<const FactPlan *>(plan)->host;
    info7[0] = h.K;
    info7[1] = h.Nr;
    info7[2] = h.Nc;
    info7[3] = h.n_lo;
    info7[4] = h.RB;
    info7[5] = h.KS;
    info7[6] = (h.ok ? 1 : 0) | (fact16_supported(h) ? 2 : 0) | (fact64_supported(h) ? 4 : 0) |
               (h.fold ? 8 : 0);
    return FEW_OK;
}

void few_mode_plan_destroy(void *plan)
{
    if (!plan) return;
    FactPlan *pl = static_cast<FactPlan *>(plan);
    if (pl->dev_buf) (void)hipFree(pl->dev_buf);
    delete pl;
}

size_t few_plan_workspace_bytes(const void *plan, int init_len, int out_len, int precision)
{
    (void)out_len;
    const int L = init_len < 2 ? 2 : init_len;
    size_t b = plan_layout(L).total;
    if (plan && (precision == FEW_PRECISION_MIXED || precision == FEW_PRECISION_MIXED16 ||
                 precision == FEW_PRECISION_FACT64)) {
        const FactPlanHost &h = static_cast<const FactPlan *>(plan)->host;
        if (h.ok) b = fact_ws_bytes(h, L, precision);
    }
    return b;
}

int few_get_waveform_plan_async(void *d_waveform, const double *interp_array,
                                const double *d_phase_interp_t, const double *phase_interp_coeffs,
                                const void *plan, const int *d_m, const int *d_k, const int *d_n,
                                int init_len, int out_len, int num_teuk_modes, const void *d_Ylms,
                                double delta_t, const double *d_h_t, int precision, int variant,
                                int stage, void *workspace, size_t workspace_bytes, void *stream)
{
    int rc = check_mode_args(d_waveform, interp_array, phase_interp_coeffs, d_m, d_k, d_n, init_len,
                             out_len, num_teuk_modes, d_Ylms, delta_t);
    if (rc) return rc;
    if (!d_phase_interp_t || !d_h_t || !workspace) {
        set_error("get_waveform_plan_async: null device time array or workspace");
        return FEW_ERR_ARG;
    }
    if (stage < 1 || stage > 7 || (stage == 4 || stage == 5)) {
        set_error("get_waveform_plan_async: stage must be 1 (prepare), 2 (sum) or 3 (both), "
                  "plus 4 (overwrite) with the sum");
        return FEW_ERR_ARG;
    }
    return run_waveform(static_cast<double2 *>(d_waveform), interp_array, d_phase_interp_t,
                        phase_interp_coeffs, d_m, d_k, d_n, init_len, out_len, num_teuk_modes,
                        static_cast<const double2 *>(d_Ylms), delta_t, d_h_t, precision,
                        static_cast<char *>(workspace), workspace_bytes,
                        reinterpret_cast<hipStream_t>(stream), static_cast<const FactPlan *>(plan),
                        variant, stage);
}

size_t few_adjoint_workspace_bytes(int init_len, int out_len, int num_modes, int precision)
{
    (void)precision;
    if (init_len < 2 || out_len < 0 || num_modes < 0) return 0;
    return adj_layout(init_len, out_len, num_modes).total;
}

int few_waveform_adjoint_async(const void *grad_wave, const double *interp_array,
                               const double *phase_interp_t, const double *phase_interp_coeffs,
                               const int *d_m, const int *d_k, const int *d_n, int init_len,
                               int out_len, int num_teuk_modes, const void *d_Ylms, double delta_t,
                               const double *h_t, double *g_interp, double *g_phase, int precision,
                               void *workspace, size_t workspace_bytes, void *stream)
{
    int rc = check_mode_args(grad_wave, interp_array, phase_interp_coeffs, d_m, d_k, d_n, init_len,
                             out_len, num_teuk_modes, d_Ylms, delta_t);
    if (rc) return rc;
    if (!phase_interp_t || !h_t || !g_interp || !g_phase || !workspace) {
        set_error("waveform_adjoint: null pointer argument");
        return FEW_ERR_ARG;
    }
    if (precision != FEW_PRECISION_FP64) {
        set_error("waveform_adjoint: precision %d is not available", precision);
        return FEW_ERR_UNSUPPORTED;
    }
    const int L = init_len, K = num_teuk_modes;
    const AdjLayout al = adj_layout(L, out_len, K);
    if (workspace_bytes < al.total) {
        set_error("waveform_adjoint: workspace of %zu bytes, need %zu", workspace_bytes, al.total);
        return FEW_ERR_ARG;
    }
    if (al.max_chunks > 0x7fffffffL || al.groups > 65535) {
        set_error("waveform_adjoint: problem too large (%d samples, %d modes)", out_len, K);
        return FEW_ERR_LIMIT;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    char *ws = static_cast<char *>(workspace);
    Segment *segs = reinterpret_cast<Segment *>(ws + al.segs);
    int *prefix = reinterpret_cast<int *>(ws + al.prefix);
    PlanArgs p{h_t, phase_interp_t, L, delta_t, out_len, ADJ_CHUNK, segs, prefix};
    rc = launch_plan(p, st);
    if (rc) return rc;
    AdjArgs a;
    a.grad = static_cast<const double2 *>(grad_wave);
    a.interp = interp_array;
    a.pcoef = phase_interp_coeffs;
    a.m = d_m;
    a.k = d_k;
    a.n = d_n;
    a.ylm = static_cast<const double2 *>(d_Ylms);
    a.segs = segs;
    a.prefix = prefix;
    a.nseg = L - 1;
    a.L = L;
    a.K = K;
    a.dt = delta_t;
    a.chunk = ADJ_CHUNK;
    a.groups = al.groups;
    a.part_amp = reinterpret_cast<double *>(ws + al.amp);
    a.part_phase = reinterpret_cast<double *>(ws + al.phase);
    if (K > 0 && out_len > 0) {
        rc = launch_adjoint_fp64(a, (int)al.max_chunks, st);
        if (rc) return rc;
    }
    AdjReduceArgs r;
    r.prefix = prefix;
    r.part_amp = a.part_amp;
    r.part_phase = a.part_phase;
    r.nseg = L - 1;
    r.L = L;
    r.K = K;
    r.groups = (K > 0 && out_len > 0) ? al.groups : 0;
    r.g_interp = g_interp;
    r.g_phase = g_phase;
    // with no samples the planner leaves every chunk count at zero and the
    // reduction writes zeros
    return launch_adjoint_reduce(r, st);
}

size_t few_plan_adjoint_workspace_bytes(const void *plan, int init_len, int out_len, int precision)
{
    if (init_len < 2 || out_len < 0) return 0;
    if ((precision != FEW_PRECISION_MIXED16 && precision != FEW_PRECISION_FACT64) || !plan)
        return adj_layout(init_len, out_len, plan ? static_cast<const FactPlan *>(plan)->host.K : 0)
            .total;
    const FactPlanHost &h = static_cast<const FactPlan *>(plan)->host;
    return (precision == FEW_PRECISION_FACT64 ? adj64_layout(h, init_len, out_len)
                                              : adj16_layout(h, init_len, out_len)).total;
}

int few_waveform_adjoint_plan_async(const void *grad_wave, const double *interp_array,
                                    const double *phase_interp_t, const double *phase_interp_coeffs,
                                    const void *plan, const int *d_m, const int *d_k, const int *d_n,
                                    int init_len, int out_len, int num_teuk_modes, const void *d_Ylms,
                                    double delta_t, const double *h_t, double *g_interp,
                                    double *g_phase, int precision, void *workspace,
                                    size_t workspace_bytes, void *stream)
{
    if (precision == FEW_PRECISION_FP64)
        return few_waveform_adjoint_async(grad_wave, interp_array, phase_interp_t,
                                          phase_interp_coeffs, d_m, d_k, d_n, init_len, out_len,
                                          num_teuk_modes, d_Ylms, delta_t, h_t, g_interp, g_phase,
                                          precision, workspace, workspace_bytes, stream);
    int rc = check_mode_args(grad_wave, interp_array, phase_interp_coeffs, d_m, d_k, d_n, init_len,
                             out_len, num_teuk_modes, d_Ylms, delta_t);
    if (rc) return rc;
    if (precision != FEW_PRECISION_MIXED16 && precision != FEW_PRECISION_FACT64) {
        set_error("waveform_adjoint: precision %d is not available", precision);
        return FEW_ERR_UNSUPPORTED;
    }
    const bool f64 = precision == FEW_PRECISION_FACT64;
    if (!plan || !phase_interp_t || !h_t || !g_interp || !g_phase || !workspace) {
        set_error("waveform_adjoint: null pointer argument");
        return FEW_ERR_ARG;
    }
    if (phase_interp_t != h_t) {
        set_error("waveform_adjoint: the mixed adjoint needs the phase and amplitude knots to be "
                  "the same array (as InterpolatedModeSum passes them)");
        return FEW_ERR_UNSUPPORTED;
    }
    const FactPlan *pl = static_cast<const FactPlan *>(plan);
    const FactPlanHost &ph = pl->host;
    if (ph.K != num_teuk_modes) {
        set_error("waveform_adjoint: plan built for %d modes, called with %d", ph.K, num_teuk_modes);
        return FEW_ERR_ARG;
    }
    const int L = init_len;
    const Adj16Layout al = f64 ? adj64_layout(ph, L, out_len) : adj16_layout(ph, L, out_len);
    if (!al.ok) {
        set_error("waveform_adjoint: mode set too large for the factorised adjoint (%d rows, %d "
                  "columns)", ph.Nr, ph.Nc);
        return FEW_ERR_UNSUPPORTED;
    }
    if (workspace_bytes < al.total) {
        set_error("waveform_adjoint: workspace of %zu bytes, need %zu", workspace_bytes, al.total);
        return FEW_ERR_ARG;
    }
    if (al.max_chunks > 0x7fffffffL) {
        set_error("waveform_adjoint: output too long (%d samples)", out_len);
        return FEW_ERR_LIMIT;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    char *ws = static_cast<char *>(workspace);
    Segment *segs = reinterpret_cast<Segment *>(ws + al.segs);
    int *prefix = reinterpret_cast<int *>(ws + al.prefix);
    const int chunk = f64 ? A64_CHUNK : A16_CHUNK;
    PlanArgs p{h_t, phase_interp_t, L, delta_t, out_len, chunk, segs, prefix};
    rc = launch_plan(p, st);
    if (rc) return rc;
    Adj16Args a;
    a.grad = static_cast<const double2 *>(grad_wave);
    a.pcoef = phase_interp_coeffs;
    a.segs = segs;
    a.prefix = prefix;
    a.nseg = L - 1;
    a.dt = delta_t;
    a.chunk = chunk;
    a.plan = pl->dev;
    a.RBM = al.RBM;
    a.NBC = al.NBC;
    a.part = f64 ? nullptr : reinterpret_cast<float *>(ws + al.part);
    a.part64 = f64 ? reinterpret_cast<double *>(ws + al.part) : nullptr;
    if (out_len > 0) {
        rc = f64 ? launch_adjoint64(a, (int)al.max_chunks, st)
                 : launch_adjoint16(a, (int)al.max_chunks, st);
        if (rc) return rc;
    }
    Adj16FinishArgs f;
    f.plan = pl->dev;
    f.M = reinterpret_cast<double *>(ws + al.M);
    f.tile = al.tile;
    f.ncol = al.NBCM * 16;
    f.h_t = h_t;
    f.L = L;
    f.interp = interp_array;
    f.ylm = static_cast<const double2 *>(d_Ylms);
    f.m = d_m;
    f.k = d_k;
    f.n = d_n;
    f.g_interp = g_interp;
    f.g_phase = g_phase;
    f.part_phase = reinterpret_cast<double *>(ws + al.phase);
    if (f64) {
        rc = launch_adjoint64_reduce(a, al.tile, f.ncol, f.M, st);
        if (rc) return rc;
        return launch_adjoint16_finish(a, f, st, false);
    }
    return launch_adjoint16_finish(a, f, st, true);
}

int few_spline_adjoint_async(const double *t, const double *g_coef, long g_sys, long g_p,
                             long g_knot, int nsys, int length, double *g_y, long gy_sys,
                             long gy_knot, double *scratch, void *stream)
{
    if (!t || !g_coef || !g_y || !scratch) {
        set_error("spline_adjoint: null pointer argument");
        return FEW_ERR_ARG;
    }
    if (nsys < 0 || length < 4) {
        set_error("spline_adjoint: need length >= 4 (the 3-knot not-a-knot system is singular) "
                  "and nsys >= 0 (got %d, %d)", length, nsys);
        return FEW_ERR_ARG;
    }
    SplineAdjArgs a;
    a.t = t;
    a.grp[0] = SplineAdjGroup{g_coef, g_sys, g_p, g_knot, g_y, gy_sys, gy_knot, nsys};
    a.ngroups = 1;
    a.L = length;
    a.ns = nsys;
    a.factor = scratch;
    a.scratch = scratch + 7L * length;
    return launch_spline_adjoint(a, reinterpret_cast<hipStream_t>(stream));
}

int few_modes_adjoint_async(const double *g_interp, const double *g_phase, const double *t,
                            int length, int num_modes, double *grad_teuk, double *grad_phase_knots,
                            double *scratch, void *stream)
{
    if (!g_interp || !t || !grad_teuk || !scratch || (g_phase && !grad_phase_knots)) {
        set_error("modes_adjoint: null pointer argument");
        return FEW_ERR_ARG;
    }
    if (num_modes <= 0 || length < 4) {
        set_error("modes_adjoint: need modes > 0 and length >= 4 (got %d, %d)", num_modes, length);
        return FEW_ERR_ARG;
    }
    const long K = num_modes, L = length;
    SplineAdjArgs a;
    a.t = t;
    // Re parts: g_interp (4, L, 2K) columns [0, K) -> grad_teuk (L, K) complex, real slots
    a.grp[0] = SplineAdjGroup{g_interp, 1, L * 2 * K, 2 * K, grad_teuk, 2, 2 * K, num_modes};
    a.grp[1] = SplineAdjGroup{g_interp + K, 1, L * 2 * K, 2 * K, grad_teuk + 1, 2, 2 * K, num_modes};
    a.grp[2] = SplineAdjGroup{g_phase, 1, 3 * L, 3, grad_phase_knots, L, 1, g_phase ? 3 : 0};
    a.ngroups = g_phase ? 3 : 2;
    a.L = length;
    a.ns = 2 * K + (g_phase ? 3 : 0);
    a.factor = scratch;
    a.scratch = scratch + 7L * length;
    return launch_spline_adjoint(a, reinterpret_cast<hipStream_t>(stream));
}

int few_spline_eval_async(const double *t, long t_sys, long t_knot, int length, int ninterps,
                          const double *coef, long c_coef, long c_sys, long c_knot,
                          const double *tnew, long tn_sys, int npts, int deriv_order, double *out,
                          int *flags, void *stream)
{
    if (!t || !coef || !tnew || !out || !flags) {
        set_error("spline_eval: null pointer argument");
        return FEW_ERR_ARG;
    }
    if (length < 2 || ninterps < 0 || npts < 0) {
        set_error("spline_eval: invalid sizes (length=%d ninterps=%d npts=%d)", length, ninterps, npts);
        return FEW_ERR_ARG;
    }
    if (deriv_order < 0 || deriv_order > 3) {
        set_error("deriv_order must be within 0 <= deriv_order <= 3.");
        return FEW_ERR_ARG;
    }
    SplineEvalArgs a{t, t_sys, t_knot, length, ninterps, coef, c_coef, c_sys, c_knot,
                     tnew, tn_sys, npts, deriv_order, out, flags};
    return launch_spline_eval(a, reinterpret_cast<hipStream_t>(stream));
}

int few_direct_mode_sum_async(const void *teuk, const void *ylms, const double *phi_phi,
                              const double *phi_theta, const double *phi_r, const int *l,
                              const int *m, const int *k, const int *n, int num_samples,
                              int num_modes, void *out, void *stream)
{
    if (!teuk || !ylms || !phi_phi || !phi_theta || !phi_r || !l || !m || !k || !n || !out) {
        set_error("direct_mode_sum: null pointer argument");
        return FEW_ERR_ARG;
    }
    if (num_samples < 0 || num_modes < 0) {
        set_error("direct_mode_sum: invalid sizes (%d samples, %d modes)", num_samples, num_modes);
        return FEW_ERR_ARG;
    }
    DenseSumArgs a{static_cast<const double2 *>(teuk), static_cast<const double2 *>(ylms), phi_phi,
                   phi_theta, phi_r, l, m, k, n, num_samples, num_modes,
                   static_cast<double2 *>(out)};
    return launch_dense_sum(a, reinterpret_cast<hipStream_t>(stream));
}

int few_fd_workspace_bytes(int num_teuk_modes, int num_segments, size_t *bytes)
{
    if (!bytes || num_teuk_modes < 0 || num_segments < 0) {
        set_error("fd_workspace_bytes: invalid arguments");
        return FEW_ERR_ARG;
    }
    *bytes = fd_workspace_bytes(num_teuk_modes, num_segments);
    return FEW_OK;
}

static int check_fd_args(const void *waveform, const double *interp_array,
                         const double *phase_interp_t, const double *phase_interp_coeffs,
                         const int *m, const int *k, const int *n, int num_teuk_modes,
                         const double *old_time_arr, int init_length, int data_length,
                         const double *frequencies, const int *seg_start_inds,
                         const int *seg_end_inds, int num_segments, const void *ylm_all,
                         int zero_index)
{
    if (!waveform || !interp_array || !phase_interp_t || !phase_interp_coeffs || !m || !k || !n ||
        !old_time_arr || !frequencies || !seg_start_inds || !seg_end_inds || !ylm_all) {
        set_error("get_waveform_generic_fd: null pointer argument");
        return FEW_ERR_ARG;
    }
    if (num_teuk_modes < 0 || init_length < 2 || data_length < 1 || num_segments < 0 ||
        num_segments > init_length - 1 || zero_index < 0 || zero_index >= data_length) {
        set_error("get_waveform_generic_fd: invalid sizes (modes %d, init_length %d, "
                  "data_length %d, segments %d, zero index %d)",
                  num_teuk_modes, init_length, data_length, num_segments, zero_index);
        return FEW_ERR_ARG;
    }
    return FEW_OK;
}

int few_get_waveform_generic_fd_async(void *waveform, const double *interp_array,
                                      const double *phase_interp_t,
                                      const double *phase_interp_coeffs, const int *m,
                                      const int *k, const int *n, int num_teuk_modes,
                                      double delta_t, const double *old_time_arr, int init_length,
                                      int data_length, const double *frequencies,
                                      const int *seg_start_inds, const int *seg_end_inds,
                                      int num_segments, const void *ylm_all, int zero_index,
                                      int include_minus_m, int separate_modes, void *workspace,
                                      size_t workspace_bytes, void *stream)
{
    (void)delta_t;  // carried by the reference's interface, unused by its kernel
    int rc = check_fd_args(waveform, interp_array, phase_interp_t, phase_interp_coeffs, m, k, n,
                           num_teuk_modes, old_time_arr, init_length, data_length, frequencies,
                           seg_start_inds, seg_end_inds, num_segments, ylm_all, zero_index);
    if (rc) return rc;
    if (!workspace) {
        set_error("get_waveform_generic_fd_async: null workspace");
        return FEW_ERR_ARG;
    }
    FdArgs a;
    a.wave = static_cast<double2 *>(waveform);
    a.interp = interp_array;
    a.phase_t = phase_interp_t;
    a.pcoef = phase_interp_coeffs;
    a.m = m;
    a.k = k;
    a.n = n;
    a.t_knots = old_time_arr;
    a.freqs = frequencies;
    a.seg_start = seg_start_inds;
    a.seg_end = seg_end_inds;
    a.ylm = static_cast<const double2 *>(ylm_all);
    a.prefix = nullptr;
    a.K = num_teuk_modes;
    a.L = init_length;
    a.nseg = num_segments;
    a.data_length = data_length;
    a.zero_index = zero_index;
    a.include_minus_m = include_minus_m && !separate_modes;
    a.separate_modes = separate_modes;
    return launch_fd_sum(a, workspace, workspace_bytes, reinterpret_cast<hipStream_t>(stream));
}

int few_get_waveform_generic_fd(void *waveform, double *interp_array, double *phase_interp_t,
                                double *phase_interp_coeffs, int *m, int *k, int *n,
                                int num_teuk_modes, double delta_t, double *old_time_arr,
                                int init_length, int data_length, double *frequencies,
                                int *seg_start_inds, int *seg_end_inds, int num_segments,
                                void *ylm_all, int zero_index, int include_minus_m,
                                int separate_modes, void *stream)
{
    int rc = check_fd_args(waveform, interp_array, phase_interp_t, phase_interp_coeffs, m, k, n,
                           num_teuk_modes, old_time_arr, init_length, data_length, frequencies,
                           seg_start_inds, seg_end_inds, num_segments, ylm_all, zero_index);
    if (rc) return rc;
    int cur = 0;
    FEW_HIP(hipGetDevice(&cur));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lk(g_ws_mu);
    const size_t need = fd_workspace_bytes(num_teuk_modes, num_segments);
    void *ws = nullptr;
    rc = library_workspace(cur, need, &ws);
    if (rc) return rc;
    rc = few_get_waveform_generic_fd_async(waveform, interp_array, phase_interp_t,
                                           phase_interp_coeffs, m, k, n, num_teuk_modes, delta_t,
                                           old_time_arr, init_length, data_length, frequencies,
                                           seg_start_inds, seg_end_inds, num_segments, ylm_all,
                                           zero_index, include_minus_m, separate_modes, ws, need,
                                           stream);
    if (rc) return rc;
    FEW_HIP(hipStreamSynchronize(s));
    return FEW_OK;
}

size_t few_interp2d_workspace_bytes(int mx) { return amp2d_workspace_bytes(mx); }

static int check_interp2d(const double *z, const double *tx, int nx, const double *ty, int ny,
                          const double *c, int kx, int ky, const double *x, int mx,
                          const double *y, int my, int num_indiv_c, int len_indiv_c)
{
    if (!z || !tx || !ty || !c || !x || !y) {
        set_error("interp2D: null pointer argument");
        return FEW_ERR_ARG;
    }
    if (mx != my) {
        set_error("interp2D: mx and my must be the same value.");
        return FEW_ERR_ARG;
    }
    if (kx < 0 || ky < 0 || kx > 5 || ky > 5) {
        set_error("interp2D: spline degrees must be within 0..5 (got %d, %d)", kx, ky);
        return FEW_ERR_ARG;
    }
    if (nx < 2 * (kx + 1) || ny < 2 * (ky + 1)) {
        set_error("interp2D: too few knots (%d, %d) for degrees (%d, %d)", nx, ny, kx, ky);
        return FEW_ERR_ARG;
    }
    if (mx < 0 || num_indiv_c < 0 ||
        (long)len_indiv_c < (long)(nx - kx - 1) * (ny - ky - 1)) {
        set_error("interp2D: invalid sizes (mx %d, sets %d, coefficients per set %d, need %ld)",
                  mx, num_indiv_c, len_indiv_c, (long)(nx - kx - 1) * (ny - ky - 1));
        return FEW_ERR_ARG;
    }
    return FEW_OK;
}

int few_interp2d_async(double *z, const double *tx, int nx, const double *ty, int ny,
                       const double *c, int kx, int ky, const double *x, int mx, const double *y,
                       int my, int num_indiv_c, int len_indiv_c, void *workspace,
                       size_t workspace_bytes, void *stream)
{
    int rc = check_interp2d(z, tx, nx, ty, ny, c, kx, ky, x, mx, y, my, num_indiv_c, len_indiv_c);
    if (rc) return rc;
    return launch_amp2d(z, tx, nx, ty, ny, c, kx, ky, x, mx, y, num_indiv_c, len_indiv_c, workspace,
                        workspace_bytes, reinterpret_cast<hipStream_t>(stream));
}

int few_interp2D_wrap(double *z, const double *tx, int nx, const double *ty, int ny, double *c,
                      int kx, int ky, const double *x, int mx, const double *y, int my,
                      int num_indiv_c, int len_indiv_c, void *stream)
{
    int rc = check_interp2d(z, tx, nx, ty, ny, c, kx, ky, x, mx, y, my, num_indiv_c, len_indiv_c);
    if (rc) return rc;
    int cur = 0;
    FEW_HIP(hipGetDevice(&cur));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lk(g_ws_mu);
    const size_t need = amp2d_workspace_bytes(mx);
    void *ws = nullptr;
    rc = library_workspace(cur, need > 0 ? need : 1, &ws);
    if (rc) return rc;
    rc = launch_amp2d(z, tx, nx, ty, ny, c, kx, ky, x, mx, y, num_indiv_c, len_indiv_c, ws, need, s);
    if (rc) return rc;
    FEW_HIP(hipStreamSynchronize(s));
    return FEW_OK;
}

size_t few_aak_workspace_bytes(int init_len) { return aak_workspace_bytes(init_len); }

static int check_aak(const void *waveform, const double *interp_array, int nmodes, int init_len,
                     int out_len, double delta_t, const double *h_t)
{
    if (!waveform || !interp_array || !h_t) {
        set_error("aak: null pointer argument");
        return FEW_ERR_ARG;
    }
    if (init_len < 2 || out_len < 0 || nmodes < 0 || !(delta_t > 0.0)) {
        set_error("aak: invalid sizes (init_len %d, out_len %d, nmodes %d, dt %g)", init_len,
                  out_len, nmodes, delta_t);
        return FEW_ERR_ARG;
    }
    return FEW_OK;
}

int few_aak_waveform_async(void *waveform, const double *interp_array, double M_phys,
                           double S_phys, double mu, double qS, double phiS, double qK,
                           double phiK, double dist, int nmodes, int mich, int init_len,
                           int out_len, double delta_t, const double *d_h_t, void *workspace,
                           size_t workspace_bytes, void *stream)
{
    int rc = check_aak(waveform, interp_array, nmodes, init_len, out_len, delta_t, d_h_t);
    if (rc) return rc;
    if (!workspace) {
        set_error("aak_waveform_async: null workspace");
        return FEW_ERR_ARG;
    }
    AakArgs a;
    a.wave = static_cast<double2 *>(waveform);
    a.interp = interp_array;
    a.h_t = d_h_t;
    a.M_phys = M_phys;
    a.S_phys = S_phys;
    a.mu = mu;
    a.qS = qS;
    a.phiS = phiS;
    a.qK = qK;
    a.phiK = phiK;
    a.dist = dist;
    a.dt = delta_t;
    a.nmodes = nmodes;
    a.mich = mich ? 1 : 0;
    a.L = init_len;
    a.out_len = out_len;
    a.plan = nullptr;
    return launch_aak(a, workspace, workspace_bytes, reinterpret_cast<hipStream_t>(stream));
}

int few_aak_get_waveform(void *waveform, double *interp_array, double M_phys, double S_phys,
                         double mu, double qS, double phiS, double qK, double phiK, double dist,
                         int nmodes, int mich, int init_len, int out_len, double delta_t,
                         double *h_t, void *stream)
{
    int rc = check_aak(waveform, interp_array, nmodes, init_len, out_len, delta_t, h_t);
    if (rc) return rc;
    int cur = 0;
    FEW_HIP(hipGetDevice(&cur));
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lk(g_ws_mu);
    const size_t plan = aak_workspace_bytes(init_len);
    const size_t need = ((plan + 255) & ~(size_t)255) + sizeof(double) * (size_t)init_len;
    void *ws = nullptr;
    rc = library_workspace(cur, need, &ws);
    if (rc) return rc;
    // h_t is a host array in the reference interface (gpuAAK.cu:536-552)
    double *d_ht = reinterpret_cast<double *>(static_cast<char *>(ws) + ((plan + 255) & ~(size_t)255));
    FEW_HIP(hipMemcpyAsync(d_ht, h_t, sizeof(double) * init_len, hipMemcpyHostToDevice, s));
    rc = few_aak_waveform_async(waveform, interp_array, M_phys, S_phys, mu, qS, phiS, qK, phiK,
                                dist, nmodes, mich, init_len, out_len, delta_t, d_ht, ws, plan,
                                stream);
    if (rc) return rc;
    FEW_HIP(hipStreamSynchronize(s));
    return FEW_OK;
}

int few_release_workspace(void)
{
    std::lock_guard<std::mutex> lk(g_ws_mu);
    for (auto &w : g_ws) {
        if (w.ptr) {
            FEW_HIP(hipFree(w.ptr));
            w.ptr = nullptr;
            w.bytes = 0;
        }
    }
    return FEW_OK;
}

}  // extern "C"

// ------------------------------------------------- allocation-free mode sum

// Workspace of few_get_waveform_plan_noalloc: one copy per GPU, part of the
// code object, so it exists once the library's kernels are loaded on a device.
// Host code never uses the name `scratch` as an address (on the host it is a
// stand-in symbol): the device address comes from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[FEW_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void *g_scratch_dev[64];  // device address of `scratch` per device ordinal

// Device address of this device's `scratch`, looked up on the first call on
// the device (with it current) and kept.
static int noalloc_scratch(void **out)
{
    int dev = 0;
    FEW_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) {
        set_error("device index %d out of range", dev);
        return FEW_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("get_waveform_plan_noalloc: hipGetSymbolAddress of the scratch array "
                      "failed: %s", hipGetErrorString(e));
            return FEW_ERR_HIP;
        }
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return FEW_OK;
}

extern "C" {

int few_get_waveform_plan_noalloc(void *d_waveform, const double *interp_array,
                                  const double *d_phase_interp_t,
                                  const double *phase_interp_coeffs, const void *plan,
                                  const int *d_m, const int *d_k, const int *d_n, int init_len,
                                  int out_len, int num_teuk_modes, const void *d_Ylms,
                                  double delta_t, const double *d_h_t, int precision, int variant,
                                  int stage, void *stream)
{
    int rc = check_mode_args(d_waveform, interp_array, phase_interp_coeffs, d_m, d_k, d_n, init_len,
                             out_len, num_teuk_modes, d_Ylms, delta_t);
    if (rc) return rc;
    const size_t need = few_plan_workspace_bytes(plan, init_len, out_len, precision);
    if (need > FEW_NOALLOC_SCRATCH_BYTES) {
        set_error("get_waveform_plan_noalloc: needs %zu bytes of workspace, more than the %u of "
                  "the library's scratch array (use few_get_waveform_plan_async with a caller "
                  "workspace)", need, (unsigned)FEW_NOALLOC_SCRATCH_BYTES);
        return FEW_ERR_LIMIT;
    }
    void *ws = nullptr;
    rc = noalloc_scratch(&ws);
    if (rc) return rc;
    return few_get_waveform_plan_async(d_waveform, interp_array, d_phase_interp_t,
                                       phase_interp_coeffs, plan, d_m, d_k, d_n, init_len, out_len,
                                       num_teuk_modes, d_Ylms, delta_t, d_h_t, precision, variant,
                                       stage, ws, FEW_NOALLOC_SCRATCH_BYTES, stream);
}

}  // extern "C"

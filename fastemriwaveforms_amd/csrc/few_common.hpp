// Shared definitions of the MI355X (gfx950) interpolated mode-summation library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define FEW_OK 0
#define FEW_ERR_ARG 1
#define FEW_ERR_HIP 2
#define FEW_ERR_SINGULAR 3
#define FEW_ERR_UNSUPPORTED 4
#define FEW_ERR_LIMIT 5

namespace few {

// Thread-local error message kept for the most recent failing call.
void set_error(const char *fmt, ...);

// Records the HIP error (if any) and returns FEW_ERR_HIP, else FEW_OK.
int check_hip(hipError_t e, const char *what);

#define FEW_HIP(call)                                   \
    do {                                                \
        hipError_t e_ = (call);                         \
        if (e_ != hipSuccess) return ::few::check_hip(e_, #call); \
    } while (0)

// hipFuncSetAttribute(kernel, MaxDynamicSharedMemorySize, bytes) on the
// current device, skipped when that kernel already allows as many bytes
// there (the attribute call costs microseconds of host time per launch).
int ensure_dynamic_lds(const void *kernel, int bytes);

// What a preparing call (interval plan, coefficient tiles) left in a caller's
// workspace: it depends on the plan, the precision, the samples per tile (the
// variant) and the sizes.  prep_stage records it when the call prepares and,
// for a call that only sums, refuses (FEW_ERR_ARG) a workspace prepared
// differently: the sum would run with a stale plan.
struct PrepKey {
    const void *ws;
    const void *plan;
    int precision, tile, L, out_len;
};
int prep_stage(bool prepares, bool sums, const PrepKey &k, const char *what);
// Zero fill of n complex128 values by a kernel (capturable in a graph).
int launch_zero_complex(double2 *w, long n, hipStream_t stream);

#define FEW_LDS(kernel, bytes)                                                  \
    do {                                                                        \
        int rc_ = ::few::ensure_dynamic_lds((const void *)(kernel), (int)(bytes)); \
        if (rc_) return rc_;                                                    \
    } while (0)

// One group of cubic-spline systems sharing a length; see spline.hip.
struct SplineGroup {
    const double *y;     // knot values: y[s*y_sys + i*y_knot]
    long y_sys, y_knot;
    const double *t;     // knot times: t[s*t_sys + i*t_knot] (t_sys = 0: shared row)
    long t_sys, t_knot;
    double *out;         // cubic: out[c*o_coef + s*o_sys + i*o_knot], c = 0..3
    long o_coef, o_sys, o_knot;
    int nsys;
    int dense;           // 1: write the 8-coefficient dense-output layout
                         //    out[c*(length-1)*nsys + j*nsys + s] instead
};

// Scratch of the sequential sweeps: four [knot][column] double arrays.
struct SplineScratch {
    double *sD, *sDU, *sDL, *sB;
    int *info;
};

__host__ __device__ void spline_system_core(const SplineGroup &G, int s, int n, long col, long ns,
                                            const SplineScratch &a);

// Launches the fused not-a-knot build for up to two groups of systems of a
// common length.  scratch holds 4*length*(nsys0+nsys1) doubles.
int launch_spline_build(const SplineGroup *groups, int ngroups, int length, double *scratch4[4],
                        int *info_out, hipStream_t stream);

// Per-interval description of the output samples it owns, shared by the
// mode-sum kernels (filled on the host or by a device planning kernel).
struct Segment {
    int start;        // first sample index (inclusive)
    int end;          // last sample index (exclusive)
    double h_t;       // amplitude-spline knot time of the interval
    double pt;        // phase dense-output segment start
    double pw;        // phase dense-output segment width
    double hw;        // amplitude knot spacing h_t[j+1] - h_t[j]
};

}  // namespace few

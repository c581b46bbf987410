// Bicubic (tensor-product B-spline) evaluation of many coefficient sets at
// many points: the amplitude interpolation interp2D
// (src/few/cutils/AmpInterp2D.cu:1-208, FITPACK fpbspl / fpbisp).
//
// The B-spline weights of a point do not depend on the coefficient set, so
// they are computed once per point (amp2d_basis_kernel: interval search and
// de Boor-Cox recurrence) instead of once per (set, point) pair; the sum
// kernel then spends 16 loads and 32 FP64 operations per (set, point), with
// one coefficient set per wave and one point per lane, so the output rows
// z[c_i * mx + i] are written coalesced.  Operation order is fpbisp's
// (sp += c * wx * wy, rows of the patch outer) with contraction off: the
// result equals the reference's arithmetic bit for bit.
#include <math.h>

#include "amp2d.hpp"
#include "few_common.hpp"

#pragma clang fp contract(off)

namespace few {

#define AMP_KMAX 5            // fpbspl's work array holds degree <= 5
#define AMP_BASIS 14          // per point: lx, ly, wx[6], wy[6]

// AmpInterp2D.cu:7-26
__host__ __device__ inline void amp_fpbspl(const double *t, int k, double x, int l, double *h)
{
    double hh[AMP_KMAX] = {0};
    h[0] = 1;
    for (int j = 1; j <= k; j++) {
        for (int i = 0; i < k; i++) hh[i] = h[i];
        h[0] = 0;
        for (int i = 0; i < j; i++) {
            const int li = l + i;
            const int lj = li - j;
            const double f = hh[i] / (t[li] - t[lj]);
            h[i] += f * (t[li] - x);
            h[i + 1] = f * (x - t[lj]);
        }
    }
}

// The interval (AmpInterp2D.cu:44-57) and weights of one coordinate.
__host__ __device__ inline int amp_interval(const double *t, int n, int k, double x, double *w)
{
    const int k1 = k + 1;
    const int nk1 = n - k1;
    const double tb = t[k1 - 1], te = t[nk1];
    double arg = x;
    if (arg < tb) arg = tb;
    if (arg > te) arg = te;
    int l = k1;
    while (!(arg < t[l] || l == nk1)) l++;
    for (int i = 0; i < 6; i++) w[i] = 0.0;
    amp_fpbspl(t, k, arg, l, w);
    return l - k1;
}

__host__ __device__ inline void amp_point_basis(const double *tx, int nx, const double *ty, int ny,
                                                int kx, int ky, double x, double y, double *b)
{
    const int lx = amp_interval(tx, nx, kx, x, b + 2);
    const int ly = amp_interval(ty, ny, ky, y, b + 8);
    b[0] = (double)lx;
    b[1] = (double)ly;
}

// fpbisp's sum (AmpInterp2D.cu:73-83) for one coefficient set.
__host__ __device__ inline double amp_sum(const double *c, int nky1, int kx, int ky, const double *b)
{
    const int lx = (int)b[0], ly = (int)b[1];
    int l1 = lx * nky1 + ly;
    double sp = 0;
    for (int i1 = 0; i1 <= kx; i1++) {
        int l2 = l1;
        for (int j1 = 0; j1 <= ky; j1++) sp += c[l2++] * b[2 + i1] * b[8 + j1];
        l1 += nky1;
    }
    return sp;
}

struct Amp2dArgs {
    double *z;
    const double *tx, *ty, *c, *x, *y;
    int nx, ny, kx, ky, mx, num_c, len_c;
    double *basis;            // (mx, AMP_BASIS)
};

__global__ __launch_bounds__(256) void amp2d_basis_kernel(Amp2dArgs a)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.mx) return;
    double b[AMP_BASIS];
    amp_point_basis(a.tx, a.nx, a.ty, a.ny, a.kx, a.ky, a.x[i], a.y[i], b);
    for (int e = 0; e < AMP_BASIS; e++) a.basis[(long)i * AMP_BASIS + e] = b[e];
}

// One coefficient set per wave (4 per workgroup), one point per lane.
template <int KX, int KY>
__global__ __launch_bounds__(256) void amp2d_sum_kernel(Amp2dArgs a)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    if (i >= a.mx) return;
    double b[AMP_BASIS];
    for (int e = 0; e < AMP_BASIS; e++) b[e] = a.basis[(long)i * AMP_BASIS + e];
    const int nky1 = a.ny - a.ky - 1;
    for (long ci = (long)blockIdx.y * 4 + wv; ci < a.num_c; ci += (long)gridDim.y * 4) {
        const double *c = a.c + ci * a.len_c;
        a.z[ci * a.mx + i] = amp_sum(c, nky1, KX, KY, b);
    }
}

__global__ __launch_bounds__(256) void amp2d_sum_any_kernel(Amp2dArgs a)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    if (i >= a.mx) return;
    double b[AMP_BASIS];
    for (int e = 0; e < AMP_BASIS; e++) b[e] = a.basis[(long)i * AMP_BASIS + e];
    const int nky1 = a.ny - a.ky - 1;
    for (long ci = (long)blockIdx.y * 4 + wv; ci < a.num_c; ci += (long)gridDim.y * 4)
        a.z[ci * a.mx + i] = amp_sum(a.c + ci * a.len_c, nky1, a.kx, a.ky, b);
}

size_t amp2d_workspace_bytes(int mx) { return sizeof(double) * AMP_BASIS * (size_t)(mx > 0 ? mx : 0); }

int launch_amp2d(double *z, const double *tx, int nx, const double *ty, int ny, const double *c,
                 int kx, int ky, const double *x, int mx, const double *y, int num_c, int len_c,
                 void *workspace, size_t workspace_bytes, hipStream_t stream)
{
    if (mx <= 0 || num_c <= 0) return FEW_OK;
    if (workspace_bytes < amp2d_workspace_bytes(mx)) {
        set_error("interp2D: workspace of %zu bytes, need %zu", workspace_bytes,
                  amp2d_workspace_bytes(mx));
        return FEW_ERR_ARG;
    }
    Amp2dArgs a{z, tx, ty, c, x, y, nx, ny, kx, ky, mx, num_c, len_c,
                static_cast<double *>(workspace)};
    hipLaunchKernelGGL(amp2d_basis_kernel, dim3((unsigned)((mx + 255) / 256)), dim3(256), 0, stream,
                       a);
    FEW_HIP(hipGetLastError());
    const long groups = ((long)num_c + 3) / 4;
    const dim3 grid((unsigned)((mx + 63) / 64), (unsigned)(groups < 65535 ? groups : 65535));
    if (kx == 3 && ky == 3)
        hipLaunchKernelGGL((amp2d_sum_kernel<3, 3>), grid, dim3(256), 0, stream, a);
    else
        hipLaunchKernelGGL(amp2d_sum_any_kernel, grid, dim3(256), 0, stream, a);
    FEW_HIP(hipGetLastError());
    return FEW_OK;
}

}  // namespace few

// Bicubic amplitude interpolation (interp2D, src/few/cutils/AmpInterp2D.cu).
#pragma once

#include "few_common.hpp"

namespace few {

// Workspace of launch_amp2d: the B-spline basis of every point.
size_t amp2d_workspace_bytes(int mx);
int launch_amp2d(double *z, const double *tx, int nx, const double *ty, int ny, const double *c,
                 int kx, int ky, const double *x, int mx, const double *y, int num_c, int len_c,
                 void *workspace, size_t workspace_bytes, hipStream_t stream);

}  // namespace few

// Frequency-domain interpolated sum (stationary-phase approximation),
// src/few/cutils/interpolate.cu:1142-1563 (get_waveform_generic_fd).
#pragma once

#include "few_common.hpp"

namespace few {

struct FdArgs {
    double2 *wave;              // (Nf,) accumulated, or (K, Nf) with separate_modes
    const double *interp;       // (4, 2K + 3, L): Re A, Im A, f_phi, f_theta, f_r
    const double *phase_t;      // (L,) knot times of the phase coefficients
    const double *pcoef;        // (8, 3, L - 1) dense-output phase coefficients
    const int *m, *k, *n;       // (K,)
    const double *t_knots;      // (L,) segment boundaries (old_time_arr)
    const double *freqs;        // (Nf,)
    const int *seg_start;       // (K, nseg) first frequency index (inclusive)
    const int *seg_end;         // (K, nseg) last frequency index (inclusive)
    const double2 *ylm;         // (2K,) Y+ then Y- (already times (-1)^l)
    const long *prefix;         // (K * nseg + 1,) exclusive prefix of in-range bins per item
    int K, L, nseg, data_length, zero_index;
    int include_minus_m, separate_modes;
};

// Workspace: the per-(mode, segment) bin counts and their prefix.
size_t fd_workspace_bytes(int K, int nseg);
int launch_fd_sum(FdArgs a, void *workspace, size_t workspace_bytes, hipStream_t stream);

}  // namespace few

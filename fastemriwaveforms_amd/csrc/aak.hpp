// Augmented Analytic Kludge summation (src/few/cutils/gpuAAK.cu).
#pragma once

#include "few_common.hpp"

namespace few {

struct AakArgs {
    double2 *wave;               // (out_len,) overwritten: hI - i hII
    const double *interp;        // (8, L - 1, 6): p, e, Y, Phi_phi, Phi_theta, Phi_r
    const double *h_t;           // (L,) knot times (device)
    double M_phys, S_phys, mu, qS, phiS, qK, phiK, dist, dt;
    int nmodes, mich, L, out_len;
    int *plan;                   // (L + 1) segment start indices, then the knot count
};

size_t aak_workspace_bytes(int init_len);
int launch_aak(AakArgs a, void *workspace, size_t workspace_bytes, hipStream_t stream);

}  // namespace few

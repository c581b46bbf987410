"""AAK summation (summation/aakwave.py:12-172) on the GPU.

:class:`AAKSummation` keeps the reference's ``sum`` signature: the sparse
trajectory's dense-output coefficients of (p, e, Y, Phi_phi, Phi_theta,
Phi_r), the source parameters, the number of harmonics and the response
choice; the sum itself is ``few_aak_waveform_async`` (csrc/aak.hip).  The 5PN
trajectory that produces the coefficients is outside this backend.
"""

from __future__ import annotations

import logging

import numpy as np
import torch

from ._lib import check, lib
from .summation import SummationBase, _device, _stream, as_device


class AAKSummation(SummationBase):
    """Augmented Analytic Kludge summation (summation/aakwave.py:12-172)."""

    def __init__(self, force_backend=None, **kwargs):
        SummationBase.__init__(self, **kwargs, force_backend=force_backend)
        self._ws = None

    @classmethod
    def supported_backends(cls):
        return cls.GPU_RECOMMENDED()

    def sum(self, tvec, m1, a, dist, m2, qS, phiS, qK, phiK, nmodes, interp_t, interp_coeffs,
            *args, mich=False, dt=10.0, integrate_backwards=False, **kwargs):
        """Fills ``self.waveform`` (hI - i hII, or h+ - i hx without ``mich``).

        interp_coeffs: (L-1, 6, 8) dense-output coefficients of (p, e, Y, Phi_phi,
        Phi_theta, Phi_r) on the knots interp_t, as the trajectory provides them."""
        fill_val = 1e-6
        log = logging.getLogger("fastemriwaveforms_amd")
        if qK < fill_val or qK > np.pi - fill_val:
            log.warning("qK is within 1e-6 of the poles. We shift this value automatically "
                        "away from poles by 1e-6.")
            qK = fill_val if qK < fill_val else np.pi - fill_val
        if qS < fill_val or qS > np.pi - fill_val:
            log.warning("qS is within 1e-6 of the poles. We shift this value automatically "
                        "away from poles by 1e-6.")
            qS = fill_val if qS < fill_val else np.pi - fill_val
        h_t = as_device(interp_t, torch.float64)
        init_len = len(h_t)
        if integrate_backwards:
            tv = as_device(tvec, torch.float64)
            offset = float(tv[-1]) - int(float(tv[-1]) / dt) * dt
            h_t = h_t - offset
        coeffs = as_device(interp_coeffs, torch.float64).clone()
        if abs(float(coeffs[0, 2, 0])) == 1.0:
            # equatorial orbits: Phi_theta follows Phi_phi (aakwave.py:144-145)
            coeffs[:, 4] = coeffs[:, 3]
        coeffs_in = coeffs.permute(2, 0, 1).contiguous()      # (8, L-1, 6)
        nbytes = int(lib().few_aak_workspace_bytes(init_len))
        if self._ws is None or self._ws.numel() < nbytes:
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=_device())
        rc = lib().few_aak_waveform_async(
            self.waveform.data_ptr(), coeffs_in.data_ptr(), float(m1), float(a), float(m2),
            float(qS), float(phiS), float(qK), float(phiK), float(dist), int(nmodes),
            int(bool(mich)), init_len, int(self.num_pts), float(dt), h_t.data_ptr(),
            self._ws.data_ptr(), self._ws.numel(), _stream())
        check(rc, "AAK summation")
        self._keep = (coeffs_in, h_t)

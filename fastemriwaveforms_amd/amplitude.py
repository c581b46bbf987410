"""Bicubic Teukolsky-amplitude interpolation (amplitude/ampinterp2d.py:29-190).

:class:`AmpInterp2D` keeps the reference's constructor and call: knots in w
and u, tensor-product B-spline coefficients per mode (Re and Im), evaluated
at arrays of (w, u) for all modes or a subset; the evaluation is the
interp2D HIP kernel of libfew_rocm.so (csrc/amp2d.hip).  The regime wrappers
that map (a, p, e, xI) to (w, u) and load the amplitude grids from data
files (AmpInterpKerrEccEq, AmpInterpSchwarzEcc) are outside this backend.
"""

from __future__ import annotations

import numpy as np
import torch

from ._lib import check, lib
from .backend import ParallelModuleBase
from .summation import _device, _stream, as_device


class AmpInterp2D(ParallelModuleBase):
    """Teukolsky amplitudes from bicubic splines (ampinterp2d.py:29-190).

    args:
        w_knots, u_knots: knots in w and u.
        coefficients: (num_modes, 2, len_indiv_c) spline coefficients, Re then Im.
        l_arr, m_arr, k_arr, n_arr: mode indices.
    """

    def __init__(self, w_knots, u_knots, coefficients, l_arr, m_arr, k_arr, n_arr,
                 force_backend=None):
        ParallelModuleBase.__init__(self, force_backend=force_backend)
        self.l_arr, self.m_arr, self.k_arr, self.n_arr = l_arr, m_arr, k_arr, n_arr
        self.coeff = as_device(coefficients, torch.float64)
        self.num_teuk_modes = self.coeff.shape[0]
        self.knots = [as_device(w_knots, torch.float64), as_device(u_knots, torch.float64)]
        self.len_indiv_c = self.coeff.shape[2]
        self._ws = None

    @classmethod
    def supported_backends(cls):
        return cls.GPU_RECOMMENDED()

    def __call__(self, w, u, *args, mode_indexes=None, **kwargs) -> torch.Tensor:
        """Complex amplitudes, shape (len(w), number of modes evaluated)."""
        w = torch.atleast_1d(as_device(w, torch.float64))
        u = torch.atleast_1d(as_device(u, torch.float64))
        if w.shape != u.shape:
            w, u = torch.broadcast_tensors(w, u)
        w = w.reshape(-1).contiguous()
        u = u.reshape(-1).contiguous()
        tw, tu = self.knots
        nmodes = self.num_teuk_modes if mode_indexes is None else len(mode_indexes)
        if w.numel() == 0:
            return torch.zeros((0, nmodes), dtype=torch.complex128, device=_device())
        if mode_indexes is None:
            c_in = self.coeff.reshape(-1)
        else:
            idx = as_device(np.asarray(mode_indexes), torch.int64)
            c_in = self.coeff[idx].reshape(-1).contiguous()
        num_indiv_c = 2 * nmodes
        mw = w.numel()
        z = torch.zeros(num_indiv_c * mw, dtype=torch.float64, device=_device())
        nbytes = int(lib().few_interp2d_workspace_bytes(mw))
        if self._ws is None or self._ws.numel() < nbytes:
            self._ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=_device())
        rc = lib().few_interp2d_async(z.data_ptr(), tw.data_ptr(), len(tw), tu.data_ptr(), len(tu),
                                      c_in.data_ptr(), 3, 3, w.data_ptr(), mw, u.data_ptr(), mw,
                                      num_indiv_c, self.len_indiv_c, self._ws.data_ptr(),
                                      self._ws.numel(), _stream())
        check(rc, "interp2D")
        z = z.reshape(num_indiv_c // 2, 2, mw).permute(2, 1, 0)
        return torch.complex(z[:, 0], z[:, 1])

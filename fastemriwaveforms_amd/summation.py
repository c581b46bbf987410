"""Native (device-resident, torch) counterparts of FEW's summation classes.

* :class:`CubicSplineInterpolant` -- summation/interpolatedmodesum.py:12-257
* :class:`SummationBase`          -- summation/base.py:5-144
* :class:`InterpolatedModeSum`    -- summation/interpolatedmodesum.py:260-375
* :class:`DirectModeSum`          -- summation/directmodesum.py:24-126
  (the frequency-domain sum is in :mod:`fastemriwaveforms_amd.fdinterp`)

Same constructor / call signatures, argument meanings and errors as the
reference.  Arrays live on the GPU as torch tensors (numpy inputs are moved
there); the spline build and the mode sum run in the HIP kernels of
libfew_rocm.so.  There is no CPU path: without the library or a GPU the
constructors raise.
"""

from __future__ import annotations

import logging

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .backend import ParallelModuleBase

YRSID_SI = 31558149.763545595  # cutils/global.h:23

PRECISIONS = {"fp64": _lib.FEW_PRECISION_FP64, "fact64": _lib.FEW_PRECISION_FACT64,
              "mixed": _lib.FEW_PRECISION_MIXED, "mixed16": _lib.FEW_PRECISION_MIXED16}


def _device() -> torch.device:
    from .pyinterp import _device as d
    return d()


def _stream() -> int:
    return int(torch.cuda.current_stream().cuda_stream)


def as_device(x, dtype=None) -> torch.Tensor:
    """numpy / torch -> contiguous tensor on the current GPU."""
    if isinstance(x, torch.Tensor):
        t = x.to(_device())
    else:
        t = torch.from_numpy(np.ascontiguousarray(x)).to(_device())
    if dtype is not None:
        t = t.to(dtype)
    return t.contiguous()


class CubicSplineInterpolant(ParallelModuleBase):
    """Batched not-a-knot cubic splines (summation/interpolatedmodesum.py:12-257).

    args:
        t: knot times, shape (length,) or (ninterps, length).
        y_all: values, shape (length,) or (ninterps, length).

    The coefficients are stored in the reference's post-build layout
    ``interp_array`` = flattened (4, length, ninterps) (interpolatedmodesum.py:99-108).
    """

    def __init__(self, t, y_all, **kwargs):
        ParallelModuleBase.__init__(self, **kwargs)
        y_all = as_device(y_all, torch.float64)
        if y_all.ndim == 1:
            y_all = y_all[None, :]
        self.degree = 3
        ninterps, length = y_all.shape
        self.ninterps = ninterps
        self.length = length
        t = as_device(t, torch.float64)
        if t.ndim == 1:
            if len(t) < 2:
                raise ValueError("t must have length greater than 2.")
            t_sys = 0
            self.t = t.expand(ninterps, length)
        elif t.ndim == 2:
            if t.shape[1] < 2:
                raise ValueError("t must have length greater than 2 along time axis.")
            t_sys = length
            self.t = t
        else:
            raise ValueError("t must be 1 or 2 dimensions.")
        if length < 3:
            # the reference's row assembly reads knots 0..2 (interpolate.cu:75-87)
            raise ValueError("the not-a-knot build needs at least 3 knots.")
        if self.t.shape[-1] != length:
            raise ValueError("t and y_all must have the same length along the time axis.")
        interp = torch.empty((4, length, ninterps), dtype=torch.float64, device=y_all.device)
        scratch = torch.empty(4 * length * ninterps, dtype=torch.float64, device=y_all.device)
        rc = lib().few_spline_build_async(
            y_all.data_ptr(), length, 1, t.data_ptr(), t_sys, 1, interp.data_ptr(),
            length * ninterps, 1, ninterps, ninterps, length, 0, scratch.data_ptr(), _stream())
        check(rc, "spline build")
        self._t_src = t
        self.interp_array = interp.reshape(-1)
        self.reshape_shape = (self.degree + 1, length, ninterps)

    @classmethod
    def supported_backends(cls):
        return cls.GPU_RECOMMENDED()

    def _coef(self, p: int) -> torch.Tensor:
        return self.interp_array.reshape(self.reshape_shape)[p].T

    @property
    def y(self):
        return self._coef(0)

    @property
    def c1(self):
        return self._coef(1)

    @property
    def c2(self):
        return self._coef(2)

    @property
    def c3(self):
        return self._coef(3)

    def __call__(self, tnew, deriv_order: int = 0):
        """Evaluate the splines (or derivatives 1-3) at tnew (interpolatedmodesum.py:161-257).

        One HIP launch (binary search + the reference's expressions, bitwise
        equal to numpy on the same coefficients).  Points outside the knot range
        get the edge values with a warning for order 0 and raise ValueError for
        derivatives, as in the reference."""
        if deriv_order not in (0, 1, 2, 3):
            raise ValueError("deriv_order must be within 0 <= deriv_order <= 3.")
        tnew = as_device(tnew, torch.float64)
        if tnew.ndim == 2:
            if tnew.shape[0] != self.ninterps:
                raise ValueError(
                    "If providing a 2D tnew array, must have same number of interpolants as was "
                    "entered during initialization.")
            tn_sys = tnew.shape[1]
        elif tnew.ndim <= 1:
            tnew = torch.atleast_1d(tnew)
            tn_sys = 0
        else:
            raise ValueError("tnew must be 1 or 2 dimensions.")
        npts = tnew.shape[-1]
        out = torch.empty((self.ninterps, npts), dtype=torch.float64, device=tnew.device)
        flags = torch.zeros(2, dtype=torch.int32, device=tnew.device)
        t_sys = 0 if self._t_src.ndim == 1 else self.length
        rc = lib().few_spline_eval_async(
            self._t_src.data_ptr(), t_sys, 1, self.length, self.ninterps,
            self.interp_array.data_ptr(), self.length * self.ninterps, 1, self.ninterps,
            tnew.data_ptr(), tn_sys, npts, deriv_order, out.data_ptr(), flags.data_ptr(),
            _stream())
        check(rc, "spline evaluation")
        if bool(flags.any()):
            if deriv_order != 0:
                raise ValueError(
                    "x points outside of the domain of the spline are not supported when taking "
                    "derivatives.")
            logging.getLogger("fastemriwaveforms_amd").warning(
                "(CubicSplineInterpolant) Warning: New t array outside bounds of input t array. "
                "These points are filled with edge values.")
        return out.squeeze()


class SummationBase(ParallelModuleBase):
    """Output sizing and dispatch of the summation modules (summation/base.py:5-144)."""

    def __init__(self, output_type: str = "td", pad_output: bool = False, odd_len: bool = False,
                 force_backend=None):
        ParallelModuleBase.__init__(self, force_backend=force_backend)
        self.pad_output = pad_output
        self.odd_len = odd_len
        if output_type not in ["td", "fd"]:
            raise ValueError(
                "{} waveform domain not available. Choices are 'td' (time domain) or 'fd' "
                "(frequency domain).".format(output_type))
        self.output_type = output_type
        if self.output_type == "fd":
            self.odd_len = True
        self.num_pts = None
        self.num_pts_pad = None
        self.waveform = None

    def sum(self, *args, **kwargs):
        raise NotImplementedError

    def __call__(self, t, *args, T: float = 1.0, dt: float = 10.0, **kwargs):
        t_first = float(t[0])
        t_last = float(t[-1])
        n_pts = int(T * YRSID_SI / dt)
        T = n_pts * dt
        if T < t_last:
            num_pts = int((T - t_first) / dt) + 1
            num_pts_pad = 0
        else:
            num_pts = int((t_last - t_first) / dt) + 1
            num_pts_pad = int((T - t_first) / dt) + 1 - num_pts if self.pad_output else 0
        self.num_pts = num_pts
        self.num_pts_pad = num_pts_pad
        self.dt = dt
        if self.odd_len and (self.num_pts + self.num_pts_pad) % 2 == 0:
            self.num_pts_pad = self.num_pts_pad + 1
        nwave = self.num_pts + self.num_pts_pad
        if self.output_type == "fd" and kwargs.get("f_arr") is not None:
            # user frequencies (summation/base.py:117-127), dt taken from them as there
            frequency = as_device(kwargs["f_arr"], torch.float64)
            dt = float(frequency.max() / 2)
            if not bool((frequency == torch.sort(frequency).values).all()):
                raise ValueError("Frequency array must be sorted in ascending order.")
            dfs = torch.diff(frequency)
            if not bool(torch.allclose(dfs, dfs[0].expand_as(dfs), rtol=1e-10)):
                raise ValueError("Frequency array must be equally spaced.")
            nwave = len(frequency)
        self.waveform = torch.zeros(nwave, dtype=torch.complex128, device=_device())
        self.sum(t, *args, dt=dt, **kwargs)
        return self.waveform


class InterpolatedModeSum(SummationBase):
    """Interpolated mode summation (summation/interpolatedmodesum.py:260-375).

    ``precision`` (constructor or ``sum`` keyword) selects the arithmetic:
    ``"fp64"`` is the reference's plain FP64 sum; ``"fact64"`` (the factorised
    sum in FP64 on the FP64 matrix cores, rel-L2 ~3e-10), ``"mixed"`` (FP64
    phases, FP32 MFMA contraction) and ``"mixed16"`` (FP64 phases, split-FP16
    MFMA with FP32 accumulation, rel-L2 ~1e-7) are the factorised sums.  They
    need a mode plan, built once per mode set and cached; a mode set whose tile
    does not fit raises :class:`FewRocmError`.
    """

    def __init__(self, force_backend=None, precision: str = "fp64", **kwargs):
        SummationBase.__init__(self, **kwargs, force_backend=force_backend)
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        self.precision = precision
        self._ws = None
        self._plans = {}

    @classmethod
    def supported_backends(cls):
        return cls.GPU_RECOMMENDED()

    def _workspace(self, L, N, K, prec) -> torch.Tensor:
        nbytes = int(lib().few_waveform_workspace_bytes(L, N, K, prec))
        if self._ws is None or self._ws.numel() < nbytes:
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=_device())
        return self._ws

    def _plan(self, m_arr, k_arr, n_arr):
        from .pipeline import ModePlan
        host = [np.ascontiguousarray(x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else x,
                                     np.int32) for x in (m_arr, k_arr, n_arr)]
        key = b"".join(h.tobytes() for h in host)
        plan = self._plans.get(key)
        if plan is None:
            plan = ModePlan(*host)
            self._plans = {key: plan}  # keep the most recent mode set only
        return plan

    def sum(self, t, teuk_modes, ylms, phase_interp_t, phase_interp_coeffs, l_arr, m_arr, k_arr,
            n_arr, *args, dt: float = 10.0, integrate_backwards: bool = False,
            precision: str | None = None, **kwargs):
        prec = PRECISIONS[precision or self.precision]
        t = as_device(t, torch.float64)
        teuk = as_device(teuk_modes, torch.complex128)
        init_len, K = teuk.shape
        L = init_len
        # amplitudes: y_all = [Re A; Im A] splined into (4, L, 2K)  (:326-337)
        interp = torch.empty((4, L, 2 * K), dtype=torch.float64, device=t.device)
        scratch = torch.empty(4 * L * 2 * K, dtype=torch.float64, device=t.device)
        rc = lib().few_prepare_modes_async(torch.view_as_real(teuk).data_ptr(), None, t.data_ptr(),
                                           L, K, interp.data_ptr(), None, scratch.data_ptr(),
                                           _stream())
        check(rc, "amplitude spline build")
        h_t = t
        pt = as_device(phase_interp_t, torch.float64)
        if integrate_backwards:
            # knots shifted to line up with t = 0 (:341-345)
            offset = h_t[-1] - torch.trunc(h_t[-1] / dt) * dt
            h_t = h_t - offset
            pt = pt - offset
        pcoef = as_device(phase_interp_coeffs, torch.float64).permute(2, 0, 1).contiguous()
        yl = as_device(ylms, torch.complex128).clone()
        l_dev = as_device(l_arr, torch.int64)
        yl[K:] *= torch.where(l_dev % 2 == 0, 1.0, -1.0).to(torch.complex128)
        m = as_device(m_arr, torch.int32)
        k = as_device(k_arr, torch.int32)
        n = as_device(n_arr, torch.int32)
        if prec == _lib.FEW_PRECISION_FP64:
            ws = self._workspace(L, self.num_pts, K, prec)
            rc = lib().few_get_waveform_async(
                self.waveform.data_ptr(), interp.data_ptr(), pt.data_ptr(), pcoef.data_ptr(),
                m.data_ptr(), k.data_ptr(), n.data_ptr(), L, self.num_pts, K, yl.data_ptr(),
                float(dt), h_t.data_ptr(), prec, ws.data_ptr(), ws.numel(), _stream())
        else:
            plan = self._plan(m_arr, k_arr, n_arr)
            nbytes = int(lib().few_plan_workspace_bytes(plan.handle, L, self.num_pts, prec))
            if self._ws is None or self._ws.numel() < nbytes:
                self._ws = torch.empty(nbytes, dtype=torch.uint8, device=_device())
            ws = self._ws
            rc = lib().few_get_waveform_plan_async(
                self.waveform.data_ptr(), interp.data_ptr(), pt.data_ptr(), pcoef.data_ptr(),
                plan.handle, m.data_ptr(), k.data_ptr(), n.data_ptr(), L, self.num_pts, K,
                yl.data_ptr(), float(dt), h_t.data_ptr(), prec, 0, 3, ws.data_ptr(), ws.numel(),
                _stream())
        check(rc, "get_waveform")
        self._keep = (interp, scratch, pcoef, yl, m, k, n, pt, h_t)


class DirectModeSum(SummationBase):
    """Direct summation over dense per-sample amplitudes and phases
    (summation/directmodesum.py:6-126), one HIP launch, FP64.

    ``sum(t, teuk_modes, ylms, phase_interp_t, phases_in, l, m, k, n)`` with
    teuk_modes (len(t), K) complex, ylms (2K,) complex (Y+ then Y-, the
    (-1)^l of the m < 0 term applied inside, as in the reference) and
    phases_in = (Phi_phi, Phi_theta, Phi_r) at the samples.
    """

    @classmethod
    def supported_backends(cls):
        return cls.GPU_RECOMMENDED()

    def sum(self, t, teuk_modes, ylms, phase_interp_t, phases_in, l_arr, m_arr, k_arr, n_arr,
            *args, dt: float = 10.0, integrate_backwards: bool = False, **kwargs):
        if integrate_backwards:
            raise NotImplementedError  # as in the reference (directmodesum.py:87-92)
        Phi_phi, Phi_theta, Phi_r = phases_in
        teuk = as_device(teuk_modes, torch.complex128)
        N, K = teuk.shape
        yl = as_device(ylms, torch.complex128)
        pp, pt, pr = (as_device(x, torch.float64) for x in (Phi_phi, Phi_theta, Phi_r))
        ints = [as_device(x, torch.int32) for x in (l_arr, m_arr, k_arr, n_arr)]
        if yl.numel() < 2 * K or min(x.numel() for x in (pp, pt, pr)) < N:
            raise ValueError("DirectModeSum: arrays smaller than teuk_modes implies")
        out = torch.empty(N, dtype=torch.complex128, device=teuk.device)
        rc = lib().few_direct_mode_sum_async(
            teuk.data_ptr(), yl.data_ptr(), pp.data_ptr(), pt.data_ptr(), pr.data_ptr(),
            ints[0].data_ptr(), ints[1].data_ptr(), ints[2].data_ptr(), ints[3].data_ptr(), N, K,
            out.data_ptr(), _stream())
        check(rc, "direct_mode_sum")
        self.waveform = out

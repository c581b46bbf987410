"""Frequency-domain interpolated mode sum (summation/fdinterp.py:92-542).

:class:`FDInterpolatedModeSum` has the reference's ``sum`` signature and
output (h+ and hx on the two-sided frequency grid, shape (2, Nf)).  The
splines, the segment-to-frequency-bin bookkeeping and the final h+/hx
assembly run on the GPU (torch for the small glue, the spline kernels of
libfew_rocm.so for the splines), and the stationary-phase sum itself is
``few_get_waveform_generic_fd_async`` (csrc/fd.hip).

The one input the reference computes that lies outside this backend is the
set of fundamental frequencies along the trajectory
(``get_fundamental_frequencies(a, p, e, xI)``, Kerr geodesics): pass them in
Hz as ``frequencies=(f_phi, f_theta, f_r)``.
"""

from __future__ import annotations

import numpy as np
import torch

from ._lib import check, lib
from .summation import CubicSplineInterpolant, SummationBase, _device, _stream, as_device


def frequency_grid(num_pts: int, num_pts_pad: int, dt: float) -> np.ndarray:
    """The default two-sided grid (fdinterp.py:412-421); computed in numpy as
    the reference does, so the bins are bit-identical."""
    Len = num_pts + num_pts_pad
    return np.hstack((np.arange(-(Len // 2), 0), np.arange(0, (Len - 1) // 2 + 1))) / (Len * dt)


def segment_frequency_inds(t, f_phi, f_theta, f_r, m, k, n, frequency: torch.Tensor,
                           f_arr_given: bool = False):
    """fdinterp.py:241-465: first/last frequency index of every (mode, segment).

    Returns (start_inds, end_inds) as (K, L-1) int32 device tensors and the
    zero index.  Mirrors the reference step by step, including its quirks:
    the turnover correction applies only when exactly one turnover exists in
    total (with more, the reference's concatenation of a (n, 2) with a (1, n)
    array raises ValueError, which it catches and ignores), and indices are
    truncated after abs()."""
    dev = _device()
    t = as_device(t, torch.float64)
    f_phi = as_device(f_phi, torch.float64)
    f_theta = as_device(f_theta, torch.float64)
    f_r = as_device(f_r, torch.float64)
    m = as_device(m, torch.int64)
    k = as_device(k, torch.int64)
    n = as_device(n, torch.int64)
    freqs_spline = CubicSplineInterpolant(t, torch.stack([f_phi, f_theta, f_r]))
    # numpy's linspace arithmetic: i * step + start, last point = stop
    t_lo, t_hi = float(t.min()), float(t.max())
    step = (t_hi - t_lo) / (5000 - 1)
    t_new = torch.arange(5000, dtype=torch.float64, device=dev) * step + t_lo
    t_new[-1] = t_hi
    new_freqs = freqs_spline(t_new)
    seg_freqs = m[:, None] * f_phi[None, :] + k[:, None] * f_theta[None, :] + n[:, None] * f_r[None, :]
    segs = torch.sort(torch.stack([seg_freqs[:, :-1], seg_freqs[:, 1:]]).permute(1, 2, 0),
                      dim=-1).values.contiguous()
    per_mode = (m[:, None] * new_freqs[0][None, :] + k[:, None] * new_freqs[1][None, :]
                + n[:, None] * new_freqs[2][None, :])
    t_of_max = t_new[per_mode.argmax(dim=-1)]
    t_of_min = t_new[per_mode.argmin(dim=-1)]
    t_last = t[-1]
    checks = [torch.nonzero((t_of_max > 0.0) & (t_of_max < t_last)).flatten(),
              torch.nonzero((t_of_min > 0.0) & (t_of_min < t_last)).flatten()]
    fix_list = []
    if len(checks[0]) > 0:
        fix_list.append(torch.searchsorted(t, t_of_max[checks[0]]) - 1)
    if len(checks[1]) > 0:
        fix_list.append(torch.searchsorted(t, t_of_min[checks[1]]) - 1)
    if fix_list:
        fix = torch.cat(fix_list)
        chk = torch.cat(checks)
        c3, c2, c1, y0 = freqs_spline.c3, freqs_spline.c2, freqs_spline.c1, freqs_spline.y
        a = m[chk] * c3[0, fix] + k[chk] * c3[1, fix] + n[chk] * c3[2, fix]
        b = m[chk] * c2[0, fix] + k[chk] * c2[1, fix] + n[chk] * c2[2, fix]
        c = m[chk] * c1[0, fix] + k[chk] * c1[1, fix] + n[chk] * c1[2, fix]
        d = m[chk] * y0[0, fix] + k[chk] * y0[1, fix] + n[chk] * y0[2, fix]
        inner = (2 * b) ** 2 - 4 * (3 * a) * c
        if bool((inner < -1e10).any()):
            raise RuntimeError("turnover discriminant below -1e10 (the reference stops here)")
        inner = torch.where(inner < 0.0, torch.zeros_like(inner), inner)
        r1 = (-(2 * b) + torch.sqrt(inner)) / (2 * (3 * a))
        r2 = (-(2 * b) - torch.sqrt(inner)) / (2 * (3 * a))
        tr1 = t[fix] + r1
        tr2 = t[fix] + r2
        keep1 = (t[fix] < tr1) & (t[fix + 1] > tr1)
        keep2 = (t[fix] < tr2) & (t[fix + 1] > tr2)
        if bool((keep1 & keep2).any()):
            raise RuntimeError("two turnover roots in one segment (the reference stops here)")
        if bool((keep1 | keep2).any()) and len(chk) == 1:
            t_fix = tr1 * keep1 + tr2 * keep2
            x_fix = t_fix - t[fix]
            f_ext = a * x_fix ** 3 + b * x_fix ** 2 + c * x_fix + d
            srt = torch.sort(torch.cat([segs[chk, fix], f_ext[None, :]], dim=-1), dim=-1).values
            segs[chk, fix] = srt[:, [0, 2]]
    nzero = int((frequency == 0.0).sum())
    if nzero != 1:
        raise ValueError("the frequency array must hold exactly one 0.0")
    ind_zero = int(torch.nonzero(frequency == 0.0)[0, 0])
    if ind_zero != int(len(frequency) / 2):
        raise ValueError("0.0 must sit at the centre of the frequency array")
    first = frequency[0]
    df = frequency[1] - frequency[0]
    inds_check = ((segs - first) / df).abs().to(torch.int64)
    if f_arr_given:
        mask_pos = segs > 0.0
        pos = ((segs - frequency[ind_zero + 1]) / df).abs().to(torch.int64) + ind_zero + 1
        inds_check = torch.where(mask_pos, pos, inds_check)
    start_inds = (inds_check[:, :, 0] + 1).to(torch.int32).contiguous()
    end_inds = inds_check[:, :, 1].to(torch.int32).contiguous()
    return start_inds, end_inds, ind_zero


def assemble_polarizations(waveform: torch.Tensor, frequency: torch.Tensor,
                           mask_positive: bool = False) -> torch.Tensor:
    """fdinterp.py:526-542: h+ and hx from the two-sided sum."""
    fd_sig = -torch.flip(waveform, dims=(-1,))
    fl = torch.flip(fd_sig, dims=(-1,))
    p = torch.complex(torch.real(fd_sig + fl) / 2.0, torch.imag(fd_sig - fl) / 2.0)
    c = torch.complex(-torch.imag(fd_sig + fl) / 2.0, torch.real(fd_sig - fl) / 2.0)
    if mask_positive:
        mask = frequency >= 0.0
        return torch.stack((p[mask], c[mask]))
    return torch.stack((p, c))


class FDInterpolatedModeSum(SummationBase):
    """Frequency-domain interpolated mode sum (summation/fdinterp.py:92-542)."""

    def __init__(self, force_backend=None, **kwargs):
        SummationBase.__init__(
            self, **{key: value for key, value in kwargs.items()
                     if key in ["output_type", "pad_output", "odd_len"]},
            force_backend=force_backend)
        self._ws = None
        self.frequency = None

    @classmethod
    def supported_backends(cls):
        return cls.GPU_RECOMMENDED()

    def sum(self, t, teuk_modes, ylms, phase_interp_t, phase_interp_coeffs, l_arr, m_arr, k_arr,
            n_arr, M=None, a=None, p=None, e=None, xI=None, *args, include_minus_m=True,
            separate_modes=False, dt=10.0, f_arr=None, mask_positive=False,
            integrate_backwards=False, frequencies=None, **kwargs):
        if frequencies is None:
            raise NotImplementedError(
                "the Kerr fundamental frequencies along the trajectory "
                "(few.utils.geodesic.get_fundamental_frequencies) are outside this backend; "
                "pass frequencies=(f_phi, f_theta, f_r) in Hz at the knots t")
        t = as_device(t, torch.float64)
        teuk = as_device(teuk_modes, torch.complex128)
        init_len, K = teuk.shape
        f_phi, f_theta, f_r = (as_device(f, torch.float64) for f in frequencies)
        # spline of [Re A; Im A; f_phi; f_theta; f_r], laid out (4, 2K+3, L)
        y_all = torch.cat([teuk.real.T, teuk.imag.T, f_phi[None], f_theta[None], f_r[None]])
        spline = CubicSplineInterpolant(t, y_all)
        spline_in = spline.interp_array.reshape(spline.reshape_shape).permute(0, 2, 1).contiguous()
        if f_arr is not None:
            self.frequency = as_device(f_arr, torch.float64)
            if len(self.frequency) == 0:
                raise ValueError("Input f_arr kwarg has zero length.")
        else:
            self.frequency = torch.from_numpy(
                frequency_grid(self.num_pts, self.num_pts_pad, dt)).to(_device())
        start_inds, end_inds, zero_index = segment_frequency_inds(
            t, f_phi, f_theta, f_r, m_arr, k_arr, n_arr, self.frequency,
            f_arr_given=f_arr is not None)
        data_length = len(self.frequency)
        if separate_modes:
            include_minus_m = False
            self.waveform = torch.zeros((K, data_length), dtype=torch.complex128, device=_device())
        pt = as_device(phase_interp_t, torch.float64)
        pcoef = as_device(phase_interp_coeffs, torch.float64).permute(2, 1, 0).contiguous()
        yl = as_device(ylms, torch.complex128).clone()
        l_dev = as_device(l_arr, torch.int64)
        yl[K:] *= torch.where(l_dev % 2 == 0, 1.0, -1.0).to(torch.complex128)
        m = as_device(m_arr, torch.int32)
        k = as_device(k_arr, torch.int32)
        n = as_device(n_arr, torch.int32)
        import ctypes
        nbytes = ctypes.c_size_t(0)
        check(lib().few_fd_workspace_bytes(K, init_len - 1, ctypes.byref(nbytes)),
              "fd workspace size")
        if self._ws is None or self._ws.numel() < nbytes.value:
            self._ws = torch.empty(max(nbytes.value, 1), dtype=torch.uint8, device=_device())
        rc = lib().few_get_waveform_generic_fd_async(
            self.waveform.data_ptr(), spline_in.data_ptr(), pt.data_ptr(), pcoef.data_ptr(),
            m.data_ptr(), k.data_ptr(), n.data_ptr(), K, float(dt), t.data_ptr(), init_len,
            data_length, self.frequency.data_ptr(), start_inds.data_ptr(), end_inds.data_ptr(),
            init_len - 1, yl.data_ptr(), zero_index, int(bool(include_minus_m)),
            int(bool(separate_modes)), self._ws.data_ptr(), self._ws.numel(), _stream())
        check(rc, "get_waveform_generic_fd")
        self.waveform = assemble_polarizations(self.waveform, self.frequency, mask_positive)

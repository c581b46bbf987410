"""Backend methods with the reference's call signatures (src/few/cutils/pyinterp.pyx:25-92).

``interpolate_arrays_wrap``, ``get_waveform_wrap`` and
``get_waveform_generic_fd_wrap`` accept exactly the positional arguments the
reference's Cython wrappers take, so FEW's ``CubicSplineInterpolant`` /
``InterpolatedModeSum`` / ``FDInterpolatedModeSum`` call them unchanged.
Array arguments may be

* torch tensors on the GPU -- passed by device pointer (no copies), or
* numpy arrays (the ``numpy``-featured drop-in backend) -- staged to the GPU,
  computed there, and written back in place, mirroring what the reference's
  CPU backend leaves in them.

Scalars pass through, like ``obj_to_ptr`` (src/few/cutils/wrappers.py:13-25).

Precision of ``get_waveform_wrap`` (the reference has one, FP64): the
environment variable ``FEW_ROCM_PRECISION`` or :func:`set_precision` selects
``"fp64"`` (default: the reference's arithmetic), ``"mixed16"`` or
``"mixed"`` (the factorised MFMA sums, rel-L2 ~1.5e-7 against FP64).  The
factorised sums need a mode plan, built once per mode set (m, k, n) and
cached; a mode set whose tiles do not fit runs in FP64.
"""

from __future__ import annotations

import logging
import os

import numpy as np
import torch

from . import _lib
from ._lib import check, lib

__all__ = ["interpolate_arrays_wrap", "get_waveform_wrap", "get_waveform_generic_fd_wrap",
           "interp2D", "pyWaveform", "set_precision", "get_precision"]

_PRECISION_CODES = {"fp64": _lib.FEW_PRECISION_FP64, "fact64": _lib.FEW_PRECISION_FACT64,
                    "mixed": _lib.FEW_PRECISION_MIXED, "mixed16": _lib.FEW_PRECISION_MIXED16}


def _check_precision(p: str) -> str:
    if p not in _PRECISION_CODES:
        raise ValueError(f"precision must be one of {sorted(_PRECISION_CODES)}, got {p!r}")
    return p


_precision = _check_precision(os.environ.get("FEW_ROCM_PRECISION", "fp64"))


def set_precision(precision: str) -> None:
    """Arithmetic of get_waveform_wrap: "fp64" (the reference's), "fact64" (the
    factorised sum in FP64, rel-L2 ~3e-10), "mixed16" or "mixed" (~1e-7)."""
    global _precision
    _precision = _check_precision(precision)


def get_precision() -> str:
    return _precision


_plans: dict = {}        # mode-set key -> ModePlan (most recent few)
_workspaces: dict = {}   # device index -> uint8 tensor


def _mode_plan(m, k, n):
    from .pipeline import ModePlan
    host = [np.ascontiguousarray(x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else x,
                                 np.int32) for x in (m, k, n)]
    key = b"".join(h.tobytes() for h in host)
    plan = _plans.get(key)
    if plan is None:
        if len(_plans) >= 4:
            _plans.pop(next(iter(_plans)))
        plan = ModePlan(*host)
        _plans[key] = plan
    return plan


def _workspace(nbytes: int) -> torch.Tensor:
    dev = _device()
    ws = _workspaces.get(dev.index)
    if ws is None or ws.numel() < nbytes:
        ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        _workspaces[dev.index] = ws
    return ws


def _device() -> torch.device:
    if not torch.cuda.is_available():
        from .exceptions import MissingHardware
        raise MissingHardware("the rocm backend needs a visible AMD GPU (torch.cuda)")
    return torch.device("cuda", torch.cuda.current_device())


def _stream_ptr() -> int:
    return int(torch.cuda.current_stream().cuda_stream)


class _Staged:
    """A device view of one array argument; writes numpy outputs back on finish()."""

    def __init__(self, obj, dtype, writeback: bool):
        self.obj = obj
        self.writeback = writeback
        if isinstance(obj, torch.Tensor):
            if not obj.is_cuda:
                self.dev = obj.to(_device()).contiguous()
            else:
                if not obj.is_contiguous():
                    raise ValueError("device tensors passed to the backend must be contiguous")
                self.dev = obj
            if self.dev.dtype != dtype:
                raise TypeError(f"expected {dtype}, got {self.dev.dtype}")
        elif isinstance(obj, np.ndarray):
            arr = np.ascontiguousarray(obj)
            self.dev = torch.from_numpy(arr).to(_device())
            if self.dev.dtype != dtype:
                raise TypeError(f"expected {dtype}, got {self.dev.dtype}")
        else:
            raise TypeError(f"unsupported array argument of type {type(obj)}")

    @property
    def ptr(self) -> int:
        return int(self.dev.data_ptr())

    def finish(self) -> None:
        if not self.writeback:
            return
        if isinstance(self.obj, np.ndarray):
            self.obj[...] = self.dev.cpu().numpy().reshape(self.obj.shape)
        elif isinstance(self.obj, torch.Tensor) and self.dev is not self.obj:
            self.obj.copy_(self.dev.reshape(self.obj.shape))


def _host_f64(obj) -> np.ndarray:
    if isinstance(obj, torch.Tensor):
        return np.ascontiguousarray(obj.detach().cpu().numpy(), dtype=np.float64)
    return np.ascontiguousarray(obj, dtype=np.float64)


def interpolate_arrays_wrap(*args, **kwargs):
    """interpolate_arrays(t_arr, interp_array, ninterps, length, B, upper_diag, diag, lower_diag).

    src/few/cutils/pyinterp.pyx:25-38.  interp_array is updated in place.
    """
    t_arr, interp_array, ninterps, length, B, upper_diag, diag, lower_diag = args
    f64 = torch.float64
    st = [_Staged(t_arr, f64, False), _Staged(interp_array, f64, True), _Staged(B, f64, True),
          _Staged(upper_diag, f64, True), _Staged(diag, f64, True), _Staged(lower_diag, f64, True)]
    n = int(ninterps) * int(length)
    for s, size in zip(st, (n, 4 * n, n, n, n, n)):
        if s.dev.numel() < size:
            raise ValueError("interpolate_arrays: array smaller than the sizes imply")
    rc = lib().few_interpolate_arrays(st[0].ptr, st[1].ptr, int(ninterps), int(length), st[2].ptr,
                                      st[3].ptr, st[4].ptr, st[5].ptr, _stream_ptr())
    check(rc, "interpolate_arrays")
    for s in st:
        s.finish()


def get_waveform_wrap(*args, **kwargs):
    """get_waveform(d_waveform, interp_array, phase_interp_t, phase_interp_coeffs, d_m, d_k, d_n,
    init_len, out_len, num_teuk_modes, d_Ylms, delta_t, h_t, dev).

    src/few/cutils/pyinterp.pyx:41-61.  d_waveform is accumulated into in place.
    phase_interp_t and h_t are host arrays in the reference; device tensors
    are accepted and copied to the host here, as the reference's Python side
    does with ``t.get()`` (summation/interpolatedmodesum.py:339).
    """
    (d_waveform, interp_array, phase_interp_t, phase_interp_coeffs, d_m, d_k, d_n, init_len,
     out_len, num_teuk_modes, d_Ylms, delta_t, h_t, dev) = args
    wave = _Staged(d_waveform, torch.complex128, True)
    interp = _Staged(interp_array, torch.float64, False)
    pcoef = _Staged(phase_interp_coeffs, torch.float64, False)
    m = _Staged(d_m, torch.int32, False)
    k = _Staged(d_k, torch.int32, False)
    n = _Staged(d_n, torch.int32, False)
    ylm = _Staged(d_Ylms, torch.complex128, False)
    pt = _host_f64(phase_interp_t)
    ht = _host_f64(h_t)
    L, N, K = int(init_len), int(out_len), int(num_teuk_modes)
    if len(pt) < L or len(ht) < L:
        raise ValueError("get_waveform: time arrays shorter than init_len")
    if (wave.dev.numel() < N or interp.dev.numel() < 4 * L * 2 * K
            or pcoef.dev.numel() < 8 * (L - 1) * 3 or min(m.dev.numel(), k.dev.numel(), n.dev.numel()) < K
            or ylm.dev.numel() < 2 * K):
        raise ValueError("get_waveform: array smaller than the sizes imply")
    # The reference selects `dev` with cudaSetDevice (interpolate.cu:753); the
    # buffers here already live on one device, which is the one used.
    dev_index = wave.dev.device.index
    prec = _precision
    plan = None
    if prec != "fp64" and K > 0:
        plan = _mode_plan(d_m, d_k, d_n)
        ok = {"mixed16": plan.supported16, "fact64": plan.supported64}.get(prec, plan.supported)
        if not ok:
            logging.getLogger("fastemriwaveforms_amd").warning(
                "get_waveform_wrap: the mode set's tiles do not fit the %s sum; running FP64", prec)
            prec = "fp64"
    if prec == "fp64":
        rc = lib().few_get_waveform(wave.ptr, interp.ptr, pt.ctypes.data, pcoef.ptr, m.ptr, k.ptr,
                                    n.ptr, L, N, K, ylm.ptr, float(delta_t), ht.ctypes.data,
                                    int(dev_index), _stream_ptr())
    else:
        # the factorised sum: device copies of the two host time arrays, the
        # cached plan and workspace; accumulates into d_waveform like the reference
        d_pt = torch.from_numpy(pt[:L]).to(wave.dev.device)
        d_ht = torch.from_numpy(ht[:L]).to(wave.dev.device)
        code = _PRECISION_CODES[prec]
        ws = _workspace(plan.workspace_bytes(L, N, prec))
        rc = lib().few_get_waveform_plan_async(
            wave.ptr, interp.ptr, d_pt.data_ptr(), pcoef.ptr, plan.handle, m.ptr, k.ptr, n.ptr, L,
            N, K, ylm.ptr, float(delta_t), d_ht.data_ptr(), code, 0, 3, ws.data_ptr(), ws.numel(),
            _stream_ptr())
        if rc == 0:
            torch.cuda.current_stream().synchronize()  # synchronous, as the reference
    check(rc, "get_waveform")
    wave.finish()


def get_waveform_generic_fd_wrap(*args, **kwargs):
    """get_waveform_generic_fd(waveform, interp_array, phase_interp_t, phase_interp_coeffs,
    m_arr_in, k_arr_in, n_arr_in, num_teuk_modes, delta_t, old_time_arr, init_length,
    data_length, frequencies, mode_start_inds, mode_end_inds, num_segments, Ylm_all,
    zero_index, include_minus_m, separate_modes).

    src/few/cutils/pyinterp.pyx:64-92.  waveform is accumulated into in place
    (overwritten per mode with separate_modes).
    """
    (waveform, interp_array, phase_interp_t, phase_interp_coeffs, m_arr_in, k_arr_in, n_arr_in,
     num_teuk_modes, delta_t, old_time_arr, init_length, data_length, frequencies,
     mode_start_inds, mode_end_inds, num_segments, Ylm_all, zero_index, include_minus_m,
     separate_modes) = args
    K, L, Nf, S = int(num_teuk_modes), int(init_length), int(data_length), int(num_segments)
    wave = _Staged(waveform, torch.complex128, True)
    interp = _Staged(interp_array, torch.float64, False)
    pt = _Staged(phase_interp_t, torch.float64, False)
    pcoef = _Staged(phase_interp_coeffs, torch.float64, False)
    m = _Staged(m_arr_in, torch.int32, False)
    k = _Staged(k_arr_in, torch.int32, False)
    n = _Staged(n_arr_in, torch.int32, False)
    tk = _Staged(old_time_arr, torch.float64, False)
    fr = _Staged(frequencies, torch.float64, False)
    s0 = _Staged(mode_start_inds, torch.int32, False)
    s1 = _Staged(mode_end_inds, torch.int32, False)
    ylm = _Staged(Ylm_all, torch.complex128, False)
    if (wave.dev.numel() < (K * Nf if separate_modes else Nf)
            or interp.dev.numel() < 4 * (2 * K + 3) * L or pt.dev.numel() < L
            or pcoef.dev.numel() < 8 * 3 * (L - 1) or min(m.dev.numel(), k.dev.numel(),
                                                          n.dev.numel()) < K
            or tk.dev.numel() < S + 1 or fr.dev.numel() < Nf
            or min(s0.dev.numel(), s1.dev.numel()) < K * S or ylm.dev.numel() < 2 * K):
        raise ValueError("get_waveform_generic_fd: array smaller than the sizes imply")
    rc = lib().few_get_waveform_generic_fd(
        wave.ptr, interp.ptr, pt.ptr, pcoef.ptr, m.ptr, k.ptr, n.ptr, K, float(delta_t), tk.ptr,
        L, Nf, fr.ptr, s0.ptr, s1.ptr, S, ylm.ptr, int(zero_index), int(bool(include_minus_m)),
        int(bool(separate_modes)), _stream_ptr())
    check(rc, "get_waveform_generic_fd")
    wave.finish()


def interp2D(z, tx, nx, ty, ny, c, kx, ky, x, mx, y, my, num_indiv_c, len_indiv_c):
    """interp2D(z, tx, nx, ty, ny, c, kx, ky, x, mx, y, my, num_indiv_c, len_indiv_c).

    src/few/cutils/pyampinterp2D.pyx:15-29 (kernel AmpInterp2D.cu:86-206).
    z (num_indiv_c * mx) is written in place; mx != my raises ValueError as
    the reference's invalid_argument does.
    """
    if int(mx) != int(my):
        raise ValueError("mx and my must be the same value.")
    zs = _Staged(z, torch.float64, True)
    st = [_Staged(a, torch.float64, False) for a in (tx, ty, c, x, y)]
    if (zs.dev.numel() < int(num_indiv_c) * int(mx) or st[0].dev.numel() < int(nx)
            or st[1].dev.numel() < int(ny) or st[2].dev.numel() < int(num_indiv_c) * int(len_indiv_c)
            or st[3].dev.numel() < int(mx) or st[4].dev.numel() < int(my)):
        raise ValueError("interp2D: array smaller than the sizes imply")
    rc = lib().few_interp2D_wrap(zs.ptr, st[0].ptr, int(nx), st[1].ptr, int(ny), st[2].ptr, int(kx),
                                 int(ky), st[3].ptr, int(mx), st[4].ptr, int(my), int(num_indiv_c),
                                 int(len_indiv_c), _stream_ptr())
    check(rc, "interp2D")
    zs.finish()


def pyWaveform(*args, **kwargs):
    """pyWaveform(waveform, interp_array, M_phys, S_phys, mu, qS, phiS, qK, phiK, dist, nmodes,
    mich, init_len, out_len, delta_t, h_t).

    src/few/cutils/gpuAAKWrap.pyx:17-33 (kernel gpuAAK.cu:181-577).  waveform
    samples of the used knot segments are overwritten in place; h_t is a host
    array in the reference (device tensors are copied to the host here).
    """
    (waveform, interp_array, M_phys, S_phys, mu, qS, phiS, qK, phiK, dist, nmodes, mich,
     init_len, out_len, delta_t, h_t) = args
    wave = _Staged(waveform, torch.complex128, True)
    interp = _Staged(interp_array, torch.float64, False)
    ht = _host_f64(h_t)
    L, N = int(init_len), int(out_len)
    if len(ht) < L or wave.dev.numel() < N or interp.dev.numel() < 8 * (L - 1) * 6:
        raise ValueError("pyWaveform: array smaller than the sizes imply")
    rc = lib().few_aak_get_waveform(wave.ptr, interp.ptr, float(M_phys), float(S_phys), float(mu),
                                    float(qS), float(phiS), float(qK), float(phiK), float(dist),
                                    int(nmodes), int(bool(mich)), L, N, float(delta_t),
                                    ht.ctypes.data, _stream_ptr())
    check(rc, "pyWaveform")
    wave.finish()

"""CPU FP64 oracle for FEW's interpolated mode summation.

TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this package, and only as the checker / CPU baseline.  The shipped
package ``fastemriwaveforms_amd`` never imports it.

``few_oracle.c`` restates the reference's CPU path (file:line citations are in
that file):

* ``interpolate_arrays``  -- src/few/cutils/interpolate.cu:284-325 (with LAPACK
  dgtsv restated; pinned against ``scipy.linalg.lapack.dgtsv`` and
  ``scipy.interpolate.CubicSpline``),
* ``get_waveform``        -- src/few/cutils/interpolate.cu:726-797 (CPU branch),
* ``waveform_samples``    -- the same per-sample arithmetic at chosen indices,
* ``find_start_inds``     -- src/few/cutils/interpolate.cu:694-723,
* ``cubic_to_dense``      -- cubic piece -> 8-coefficient dense-output form
  consumed at src/few/cutils/interpolate.cu:642.

Python-level helpers restating the reference's Python glue live here too:
``spline_eval`` (summation/interpolatedmodesum.py:161-257) and
``num_points`` (summation/base.py:83-102).
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfew_oracle.so")
_lib = None

YRSID_SI = 31558149.763545595  # src/few/cutils/global.h:23


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        d = ctypes.POINTER(ctypes.c_double)
        i = ctypes.POINTER(ctypes.c_int)
        _lib.orc_dgtsv.argtypes = [ctypes.c_int, d, d, d, d]
        _lib.orc_dgtsv.restype = ctypes.c_int
        _lib.orc_interpolate_arrays.argtypes = [d, d, ctypes.c_int, ctypes.c_int, d, d, d, d]
        _lib.orc_interpolate_arrays.restype = ctypes.c_int
        _lib.orc_find_start_inds.argtypes = [i, d, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        _lib.orc_find_start_inds.restype = ctypes.c_int
        wave = [d, d, d, i, i, i, ctypes.c_int, ctypes.c_int, ctypes.c_int, d,
                ctypes.c_double, d, ctypes.c_int]
        _lib.orc_get_waveform.argtypes = [d] + wave
        _lib.orc_get_waveform.restype = ctypes.c_int
        _lib.orc_waveform_samples.argtypes = [d, ctypes.POINTER(ctypes.c_int64), ctypes.c_long] + wave
        _lib.orc_waveform_samples.restype = ctypes.c_int
        _lib.orc_cubic_to_dense.argtypes = [d, d, ctypes.c_int, ctypes.c_int, d]
        _lib.orc_cubic_to_dense.restype = None
        _lib.orc_waveform_adjoint.argtypes = [d] + wave[:-1] + [d, d, ctypes.c_int]
        _lib.orc_waveform_adjoint.restype = ctypes.c_int
        _lib.orc_spline_adjoint.argtypes = [d, d, ctypes.c_int, ctypes.c_int, d]
        _lib.orc_spline_adjoint.restype = ctypes.c_int
        _lib.orc_max_threads.argtypes = []
        _lib.orc_max_threads.restype = ctypes.c_int
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def dgtsv(dl, d, du, b):
    """Restated LAPACK dgtsv (NRHS=1). Returns (x, info); inputs untouched."""
    dl, d, du, b = (_f64(np.array(v, dtype=np.float64, copy=True)) for v in (dl, d, du, b))
    info = lib().orc_dgtsv(len(d), _dp(dl), _dp(d), _dp(du), _dp(b))
    return b, info


def interpolate_arrays(t, y_all):
    """Spline build with the reference's layout.

    t: (length,) or (ninterps, length); y_all: (ninterps, length).
    Returns (interp_array (4, ninterps, length), B slopes (ninterps, length), info).
    """
    y_all = np.atleast_2d(_f64(y_all))
    ninterps, length = y_all.shape
    t = _f64(t)
    if t.ndim == 1:
        t = np.tile(t, (ninterps, 1))
    t = _f64(t)
    interp = np.zeros((4, ninterps, length))
    interp[0] = y_all
    B = np.zeros((ninterps, length))
    ud, diag, ld = np.zeros_like(B), np.zeros_like(B), np.zeros_like(B)
    info = lib().orc_interpolate_arrays(_dp(t), _dp(interp), ninterps, length,
                                        _dp(B), _dp(ud), _dp(diag), _dp(ld))
    return interp, B, info


def find_start_inds(h_t, delta_t, out_len):
    h_t = _f64(h_t)
    start = np.zeros(len(h_t) + 1, dtype=np.int32)
    used = lib().orc_find_start_inds(_ip(start), _dp(h_t), float(delta_t), len(h_t), int(out_len))
    return start[:used], used


def cubic_to_dense(coeffs, t):
    """coeffs (4, nphase, length) -> dense-output coefficients (8, length-1, nphase)."""
    coeffs = _f64(coeffs)
    t = _f64(t)
    _, nphase, length = coeffs.shape
    out = np.zeros((8, length - 1, nphase))
    lib().orc_cubic_to_dense(_dp(coeffs), _dp(t), nphase, length, _dp(out))
    return out


def _wave_args(interp_array_tl, phase_t, phase_coeffs_flat, m, k, n, init_len, out_len, K, ylms,
               dt, h_t, nthreads):
    ylms = np.ascontiguousarray(ylms, dtype=np.complex128).view(np.float64)
    return ([_dp(interp_array_tl), _dp(phase_t), _dp(phase_coeffs_flat), _ip(m), _ip(k), _ip(n),
             int(init_len), int(out_len), int(K), _dp(ylms), float(dt), _dp(h_t), int(nthreads)],
            ylms)


def get_waveform(interp_array_tl, phase_t, phase_coeffs_flat, m, k, n, init_len, out_len,
                 num_modes, ylms, dt, h_t, nthreads=1, waveform=None):
    """Reference get_waveform (CPU branch) into a fresh (or given) complex128 buffer.

    interp_array_tl is the (4, init_len, 2K) layout produced by
    CubicSplineInterpolant (interpolatedmodesum.py:99-101); phase_coeffs_flat
    is the (8, init_len-1, 3) layout (interpolatedmodesum.py:352-354).
    """
    if waveform is None:
        waveform = np.zeros(out_len, dtype=np.complex128)
    wv = waveform.view(np.float64)
    keep = [_f64(interp_array_tl), _f64(phase_t), _f64(phase_coeffs_flat), _i32(m), _i32(k),
            _i32(n), _f64(h_t)]
    args, yl = _wave_args(keep[0], keep[1], keep[2], keep[3], keep[4], keep[5], init_len,
                          out_len, num_modes, ylms, dt, keep[6], nthreads)
    rc = lib().orc_get_waveform(_dp(wv), *args)
    if rc != 0:
        raise RuntimeError(f"oracle get_waveform failed: {rc}")
    return waveform


def waveform_samples(idx, interp_array_tl, phase_t, phase_coeffs_flat, m, k, n, init_len,
                     out_len, num_modes, ylms, dt, h_t, nthreads=0):
    """Values get_waveform would leave in a zeroed buffer, at sample indices idx."""
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    out = np.zeros(len(idx), dtype=np.complex128)
    keep = [_f64(interp_array_tl), _f64(phase_t), _f64(phase_coeffs_flat), _i32(m), _i32(k),
            _i32(n), _f64(h_t)]
    if nthreads <= 0:
        nthreads = min(8, os.cpu_count() or 1)
    args, yl = _wave_args(keep[0], keep[1], keep[2], keep[3], keep[4], keep[5], init_len,
                          out_len, num_modes, ylms, dt, keep[6], nthreads)
    rc = lib().orc_waveform_samples(_dp(out.view(np.float64)),
                                    idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                    len(idx), *args)
    if rc != 0:
        raise RuntimeError(f"oracle waveform_samples failed: {rc}")
    return out


def waveform_adjoint(grad_wave, interp_array_tl, phase_t, phase_coeffs_flat, m, k, n, init_len,
                     out_len, num_modes, ylms, dt, h_t, nthreads=0):
    """Adjoint of get_waveform for L = sum Re(conj(g) h), g = grad_wave (complex, out_len).

    Returns (g_interp (4, init_len, 2K) in the interp_array layout,
    g_phase (4, init_len, 3): dL/dc_p of phase q (cubic coefficients) on interval j).
    """
    gw = np.ascontiguousarray(grad_wave, dtype=np.complex128)
    assert len(gw) >= out_len
    keep = [_f64(interp_array_tl), _f64(phase_t), _f64(phase_coeffs_flat), _i32(m), _i32(k),
            _i32(n), _f64(h_t)]
    if nthreads <= 0:
        nthreads = min(8, os.cpu_count() or 1)
    args, yl = _wave_args(keep[0], keep[1], keep[2], keep[3], keep[4], keep[5], init_len,
                          out_len, num_modes, ylms, dt, keep[6], nthreads)
    g_interp = np.zeros((4, init_len, 2 * num_modes))
    g_phase = np.zeros((4, init_len, 3))
    rc = lib().orc_waveform_adjoint(_dp(gw.view(np.float64)), *args[:-1], _dp(g_interp),
                                    _dp(g_phase), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle waveform_adjoint failed: {rc}")
    return g_interp, g_phase


def spline_adjoint(t, g_coef):
    """Transpose of interpolate_arrays' map y -> interp_array for fixed knots t.

    g_coef: (4, ninterps, length).  Returns (g_y (ninterps, length), info)."""
    g_coef = _f64(g_coef)
    _, ninterps, length = g_coef.shape
    t = _f64(t)
    if t.ndim == 1:
        t = _f64(np.tile(t, (ninterps, 1)))
    g_y = np.zeros((ninterps, length))
    info = lib().orc_spline_adjoint(_dp(t), _dp(g_coef), ninterps, length, _dp(g_y))
    return g_y, info


def max_threads() -> int:
    return int(lib().orc_max_threads())


# ---------------------------------------------------------------------------
# Python-level restatements of the reference's Python glue.
# ---------------------------------------------------------------------------

def num_points(t_first, t_last, T_years, dt, pad_output=False, odd_len=False):
    """summation/base.py:83-108: (num_pts, num_pts_pad) of the output buffer."""
    n_pts = int(T_years * YRSID_SI / dt)
    T = n_pts * dt
    if T < t_last:
        num_pts = int((T - t_first) / dt) + 1
        num_pts_pad = 0
    else:
        num_pts = int((t_last - t_first) / dt) + 1
        num_pts_pad = int((T - t_first) / dt) + 1 - num_pts if pad_output else 0
    if odd_len and (num_pts + num_pts_pad) % 2 == 0:
        num_pts_pad += 1
    return num_pts, num_pts_pad


def spline_eval(t_knots, interp, tnew, deriv_order=0):
    """summation/interpolatedmodesum.py:119-257 on the (4, ninterps, length) layout.

    t_knots: (ninterps, length); tnew: (Nnew,) or (ninterps, Nnew).
    Edge values fill points outside the knot range (order 0 only).
    """
    t_knots = np.atleast_2d(t_knots)
    ninterps, length = interp.shape[1:]
    tnew = np.atleast_1d(np.asarray(tnew, dtype=np.float64))
    if tnew.ndim == 1:
        tnew = np.tile(tnew, (ninterps, 1))
    inds = np.zeros(tnew.shape, dtype=int)
    for i in range(ninterps):
        inds[i] = np.searchsorted(t_knots[i], tnew[i], side="right") - 1
        inds[i][tnew[i] == t_knots[i][-1]] = length - 2
    bad_left = tnew < t_knots[:, :1]
    bad_right = tnew > t_knots[:, -1:]
    inds_c = np.clip(inds, 0, length - 1)
    rows = np.arange(ninterps)[:, None]
    x = tnew - t_knots[rows, inds_c]
    y, c1, c2, c3 = (interp[p][rows, inds_c] for p in range(4))
    if deriv_order == 0:
        out = y + c1 * x + c2 * (x * x) + c3 * (x * x * x)
        out = np.where(bad_left, interp[0][:, :1], out)
        out = np.where(bad_right, interp[0][:, -1:], out)
    else:
        if bad_left.any() or bad_right.any():
            raise ValueError("derivatives outside the spline domain are not supported")
        if deriv_order == 1:
            out = c1 + 2 * c2 * x + 3 * c3 * (x * x)
        elif deriv_order == 2:
            out = 2 * c2 + 6 * c3 * x
        elif deriv_order == 3:
            out = 6 * c3
        else:
            raise ValueError("deriv_order must be within 0 <= deriv_order <= 3.")
    return out.squeeze()


def mismatch(a, b):
    """utils/utility.py:16-97: 1 - Re<a,b>/sqrt(<a,a><b,b>) over FFTs."""
    n = min(len(a), len(b))
    fa = np.fft.fft(a[:n])
    fb = np.fft.fft(b[:n])
    ac = np.dot(fa.conj(), fb) / np.sqrt(np.dot(fa.conj(), fa) * np.dot(fb.conj(), fb))
    return 1.0 - ac.real


def rel_l2(h, ref):
    h = np.asarray(h)
    ref = np.asarray(ref)
    return float(np.linalg.norm(h - ref) / np.linalg.norm(ref))

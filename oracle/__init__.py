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

``few_oracle_fd.c`` restates the frequency-domain sum:

* ``fd_waveform``         -- src/few/cutils/interpolate.cu:1142-1563 (CPU branch),
* ``fd_spa_func`` / ``fd_cube_roots`` -- its building blocks
  (interpolate.cu:1041-1090, 1113-1134), pinned against scipy's
  ``kv(1/3, .)`` and ``numpy.roots``.

``few_oracle_amp.c`` restates the bicubic amplitude interpolation:

* ``interp2d``            -- src/few/cutils/AmpInterp2D.cu:1-208 (CPU branch,
  FITPACK fpbspl / fpbisp), pinned against scipy's FITPACK ``bisplev``.

``few_oracle_aak.c`` restates the AAK summation:

* ``aak_waveform``        -- src/few/cutils/gpuAAK.cu:181-577 (CPU branch), with
  ``bessel_jn`` (Numerical Recipes, gpuAAK.cu:8-118, pinned against scipy's
  ``jv``) and ``aak_rot_coeff`` (gpuAAK.cu:136-176, pinned by vector
  identities); the circular limit is pinned by the quadrupole closed form.

``fd_segment_inds`` and ``fd_post`` restate fdinterp.py's Python glue
(summation/fdinterp.py:202-465 and 514-542) in numpy.

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
        _lib.orc_fd_waveform.argtypes = [d, d, d, d, i, i, i, ctypes.c_int, d, ctypes.c_int,
                                         ctypes.c_int, d, i, i, ctypes.c_int, d, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int]
        _lib.orc_fd_waveform.restype = ctypes.c_int
        _lib.orc_fd_spa_func.argtypes = [d, ctypes.c_long, d]
        _lib.orc_fd_spa_func.restype = None
        _lib.orc_fd_cube_roots.argtypes = [d, ctypes.c_long, d]
        _lib.orc_fd_cube_roots.restype = None
        _lib.orc_interp2d.argtypes = [d, d, ctypes.c_int, d, ctypes.c_int, d, ctypes.c_int,
                                      ctypes.c_int, d, ctypes.c_int, d, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int]
        _lib.orc_interp2d.restype = ctypes.c_int
        _lib.orc_bessel_jn.argtypes = [ctypes.c_int, ctypes.c_double]
        _lib.orc_bessel_jn.restype = ctypes.c_double
        _lib.orc_bessel_j1_export.argtypes = [ctypes.c_double]
        _lib.orc_bessel_j1_export.restype = ctypes.c_double
        _lib.orc_aak_harmonic.argtypes = [ctypes.c_int, ctypes.c_double, d, d, d]
        _lib.orc_aak_harmonic.restype = None
        _lib.orc_aak_rot_coeff.argtypes = [d, d, d, d, d, d] + [ctypes.c_double] * 6
        _lib.orc_aak_rot_coeff.restype = None
        _lib.orc_aak_start_inds.argtypes = [i, d, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        _lib.orc_aak_start_inds.restype = ctypes.c_int
        _lib.orc_aak_waveform.argtypes = [d, d] + [ctypes.c_double] * 8 + [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, d]
        _lib.orc_aak_waveform.restype = ctypes.c_int
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


def bessel_jn(n, x):
    """Numerical Recipes J_n (gpuAAK.cu:62-118); n = -1 gives gpuAAK's -J1."""
    if n == -1:
        return -lib().orc_bessel_j1_export(float(x))
    return lib().orc_bessel_jn(int(n), float(x))


def aak_harmonic(n, e):
    """gpuAAK.cu:421-446: the Bessel brackets (X, Y, J_n) of harmonic n at
    eccentricity e, as the waveform uses them."""
    out = [np.zeros(1) for _ in range(3)]
    lib().orc_aak_harmonic(int(n), float(e), *(_dp(o) for o in out))
    return tuple(float(o[0]) for o in out)


def aak_rot_coeff(iota, theta_S, phi_S, theta_K, phi_K, alpha):
    """gpuAAK.cu:136-176: (rot (4,), n, L, S, nxL, nxS)."""
    out = [np.zeros(4)] + [np.zeros(3) for _ in range(5)]
    lib().orc_aak_rot_coeff(*(_dp(o) for o in out), float(iota), float(theta_S), float(phi_S),
                            float(theta_K), float(phi_K), float(alpha))
    return out


def aak_start_inds(h_t, delta_t, out_len):
    """gpuAAK.cu:458-483: (start indices, number of knots used)."""
    h_t = _f64(h_t)
    st = np.zeros(len(h_t) + 1, dtype=np.int32)
    npts = lib().orc_aak_start_inds(_ip(st), _dp(h_t), float(delta_t), len(h_t), int(out_len))
    return st[:npts], npts


def aak_waveform(interp_coeffs, M, S, mu, qS, phiS, qK, phiK, dist, nmodes, mich, h_t, out_len,
                 delta_t):
    """get_waveform of gpuAAK.cu (CPU branch).  interp_coeffs: (8, L-1, 6) dense-output
    coefficients of (p, e, Y, Phi_phi, Phi_theta, Phi_r).  Returns (out_len,) complex."""
    interp = _f64(interp_coeffs)
    wave = np.zeros(out_len, dtype=np.complex128)
    h_t = _f64(h_t)
    rc = lib().orc_aak_waveform(_dp(wave.view(np.float64)), _dp(interp), M, S, mu, qS, phiS, qK,
                                phiK, dist, int(nmodes), int(bool(mich)), len(h_t), int(out_len),
                                float(delta_t), _dp(h_t))
    if rc != 0:
        raise RuntimeError(f"oracle aak_waveform failed: {rc}")
    return wave


def interp2d(tx, ty, c, x, y, kx=3, ky=3):
    """interp2D (AmpInterp2D.cu:86-162, CPU branch): c (num_sets, len_per_set)
    tensor-product B-spline coefficients on knots tx, ty; returns z
    (num_sets, len(x))."""
    tx, ty, x, y = _f64(tx), _f64(ty), _f64(x), _f64(y)
    c = _f64(np.atleast_2d(c))
    num, length = c.shape
    z = np.zeros((num, len(x)))
    rc = lib().orc_interp2d(_dp(z), _dp(tx), len(tx), _dp(ty), len(ty), _dp(c), kx, ky, _dp(x),
                            len(x), _dp(y), len(y), num, length)
    if rc != 0:
        raise ValueError("mx and my must be the same value.")
    return z


def fd_waveform(interp_fd, phase_t, phase_coeffs_fd, m, k, n, t, frequencies, start_inds,
                end_inds, ylms, zero_index, include_minus_m=True, separate_modes=False):
    """get_waveform_generic_fd's CPU branch (interpolate.cu:1142-1563).

    interp_fd: (4, 2K+3, L); phase_coeffs_fd: (8, 3, L-1); start/end_inds:
    (K, L-1) inclusive frequency-index ranges; ylms: (2K,) complex (the -m
    half already multiplied by (-1)^l).  Returns the complex waveform, (Nf,)
    or (K, Nf) with separate_modes."""
    interp_fd = _f64(interp_fd)
    _, nint, L = interp_fd.shape
    K = (nint - 3) // 2
    Nf = len(frequencies)
    wave = np.zeros((K, Nf) if separate_modes else (Nf,), dtype=np.complex128)
    ylm = np.ascontiguousarray(np.asarray(ylms, dtype=np.complex128))
    rc = lib().orc_fd_waveform(
        _dp(wave.view(np.float64)), _dp(interp_fd), _dp(_f64(phase_t)), _dp(_f64(phase_coeffs_fd)),
        _ip(_i32(m)), _ip(_i32(k)), _ip(_i32(n)), K, _dp(_f64(t)), L, Nf, _dp(_f64(frequencies)),
        _ip(_i32(start_inds)), _ip(_i32(end_inds)), L - 1, _dp(ylm.view(np.float64)),
        int(zero_index), int(bool(include_minus_m)), int(bool(separate_modes)))
    if rc != 0:
        raise RuntimeError(f"oracle fd_waveform failed: {rc}")
    return wave


def fd_spa_func(x):
    """SPAFunc (interpolate.cu:1041-1090) at each x."""
    x = _f64(np.atleast_1d(x))
    out = np.zeros(len(x), dtype=np.complex128)
    lib().orc_fd_spa_func(_dp(x), len(x), _dp(out.view(np.float64)))
    return out


def fd_cube_roots(abcd):
    """cube_roots (interpolate.cu:1113-1134) of rows (a, b, c, d) -> (n, 3) complex."""
    abcd = _f64(np.atleast_2d(abcd))
    out = np.zeros((len(abcd), 3), dtype=np.complex128)
    lib().orc_fd_cube_roots(_dp(abcd), len(abcd), _dp(out.view(np.float64)))
    return out


def fd_frequency_grid(num_pts, num_pts_pad, dt):
    """summation/fdinterp.py:412-421 -- the default frequency array."""
    Len = num_pts + num_pts_pad
    return np.hstack((np.arange(-(Len // 2), 0), np.arange(0, (Len - 1) // 2 + 1))) / (Len * dt)


def fd_segment_inds(t, f_phi, f_theta, f_r, m, k, n, frequency, f_arr_given=False):
    """summation/fdinterp.py:241-465 in numpy: (start_inds, end_inds), each
    (K, L-1) int, plus the zero index.

    Follows the reference step by step, including its quirks: the turnover
    correction is applied only when exactly one turnover exists in total (with
    more, concatenating a (n, 2) with a (1, n) array raises ValueError, which
    the reference catches and ignores), and the indices are truncated with
    astype(int) after abs()."""
    t = _f64(t)
    m = np.asarray(m, dtype=np.int64)
    k = np.asarray(k, dtype=np.int64)
    n = np.asarray(n, dtype=np.int64)
    y_freqs = np.asarray([f_phi, f_theta, f_r], dtype=np.float64)
    fs, _, _ = interpolate_arrays(t, y_freqs)
    t_new = np.linspace(t.min(), t.max(), 5000)
    new_freqs = spline_eval(np.tile(t, (3, 1)), fs, t_new)
    seg_freqs = m[:, None] * f_phi[None, :] + k[:, None] * f_theta[None, :] + n[:, None] * f_r[None, :]
    segs = np.sort(np.asarray([seg_freqs[:, :-1], seg_freqs[:, 1:]]).transpose(1, 2, 0), axis=-1)
    new_freqs_per_mode = (m[:, None] * new_freqs[0][None, :] + k[:, None] * new_freqs[1][None, :]
                          + n[:, None] * new_freqs[2][None, :])
    t_of_max = t_new[new_freqs_per_mode.argmax(axis=-1)]
    t_of_min = t_new[new_freqs_per_mode.argmin(axis=-1)]
    checks = [np.where((t_of_max > 0.0) & (t_of_max < t[-1]))[0],
              np.where((t_of_min > 0.0) & (t_of_min < t[-1]))[0]]
    fix_list = []
    if len(checks[0]) > 0:
        fix_list.append(np.searchsorted(t, t_of_max[checks[0]]) - 1)
    if len(checks[1]) > 0:
        fix_list.append(np.searchsorted(t, t_of_min[checks[1]]) - 1)
    c3, c2, c1, y0 = fs[3], fs[2], fs[1], fs[0]  # (3, L) each
    try:
        fix = np.concatenate(fix_list)
        chk = np.concatenate(checks)
        a = m[chk] * c3[0, fix] + k[chk] * c3[1, fix] + n[chk] * c3[2, fix]
        b = m[chk] * c2[0, fix] + k[chk] * c2[1, fix] + n[chk] * c2[2, fix]
        c = m[chk] * c1[0, fix] + k[chk] * c1[1, fix] + n[chk] * c1[2, fix]
        d = m[chk] * y0[0, fix] + k[chk] * y0[1, fix] + n[chk] * y0[2, fix]
        inner = (2 * b) ** 2 - 4 * (3 * a) * c
        if np.any(inner < -1e10):
            raise RuntimeError("turnover discriminant below -1e10 (the reference stops here)")
        inner[inner < 0.0] = 0.0
        r1 = (-(2 * b) + np.sqrt(inner)) / (2 * (3 * a))
        r2 = (-(2 * b) - np.sqrt(inner)) / (2 * (3 * a))
        tr1 = t[fix] + r1
        tr2 = t[fix] + r2
        keep1 = (t[fix] < tr1) & (t[fix + 1] > tr1)
        keep2 = (t[fix] < tr2) & (t[fix + 1] > tr2)
        if np.any(keep1 & keep2):
            raise RuntimeError("two turnover roots in one segment (the reference stops here)")
        elif not np.any(keep1 | keep2):
            pass
        else:
            t_fix = tr1 * keep1 + tr2 * keep2
            x_fix = t_fix - t[fix]
            f_ext = a * x_fix**3 + b * x_fix**2 + c * x_fix + d
            srt = np.sort(np.concatenate([segs[chk, fix], np.array([f_ext])], axis=-1), axis=-1)
            segs[chk, fix] = srt[:, np.array([0, 2])]
    except ValueError:
        pass
    assert 1 == np.sum(frequency == 0.0)
    ind_zero = np.where(frequency == 0)[0][0]
    assert ind_zero == int(len(frequency) / 2)
    first = frequency[0]
    df = frequency[1] - frequency[0]
    inds_check = np.abs((segs - first) / df).astype(int)
    if f_arr_given:
        mask_pos = segs > 0.0
        pos = np.abs((segs - frequency[ind_zero + 1]) / df).astype(int) + ind_zero + 1
        inds_check[mask_pos] = pos[mask_pos]
    start_inds = (inds_check[:, :, 0].copy() + 1).astype(int)
    end_inds = inds_check[:, :, 1].copy().astype(int)
    zero_index = int(np.where(frequency == 0.0)[0][0])
    return start_inds, end_inds, zero_index


def fd_post(waveform, frequency, mask_positive=False):
    """summation/fdinterp.py:526-542 -- h+ and hx from the two-sided sum."""
    fd_sig = -np.flip(waveform)
    p = np.real(fd_sig + np.flip(fd_sig)) / 2.0 + 1j * np.imag(fd_sig - np.flip(fd_sig)) / 2.0
    c = -np.imag(fd_sig + np.flip(fd_sig)) / 2.0 + 1j * np.real(fd_sig - np.flip(fd_sig)) / 2.0
    if mask_positive:
        mask = frequency >= 0.0
        return np.vstack((p[mask], c[mask]))
    return np.vstack((p, c))


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

"""Shared test helpers: the oracle pipeline for a synthetic input."""

import numpy as np

import oracle


def oracle_coefficients(inp):
    """Spline coefficients in the layouts get_waveform reads, via the oracle.

    Amplitudes: y_all = [Re A; Im A] (interpolatedmodesum.py:326-335) splined
    and transposed to (4, L, 2K) (:99-101).  Phases: knot values splined and
    converted to the (8, L-1, 3) dense-output layout.
    """
    y_all = np.concatenate([inp.teuk.T.real, inp.teuk.T.imag])
    interp, _, info = oracle.interpolate_arrays(inp.t, y_all)
    assert info == 0
    interp_tl = np.ascontiguousarray(np.transpose(interp, (0, 2, 1)))
    pc, _, info = oracle.interpolate_arrays(inp.t, inp.phase_knots)
    assert info == 0
    pcoef = oracle.cubic_to_dense(pc, inp.t)
    return interp_tl, pcoef


def oracle_waveform(inp, idx=None, nthreads=8):
    interp_tl, pcoef = oracle_coefficients(inp)
    args = (interp_tl, inp.t, pcoef, inp.m, inp.k, inp.n, inp.L, inp.num_pts, inp.K, inp.ylms,
            inp.dt, inp.t)
    if idx is None:
        return oracle.get_waveform(*args, nthreads=nthreads)
    return oracle.waveform_samples(idx, *args, nthreads=nthreads)


def spread_indices(N, L_bounds=None, count=2000, seed=0):
    """Sample indices spread over [0, N): ends, a uniform grid and random draws."""
    rng = np.random.default_rng(seed)
    parts = [np.arange(min(20, N)), np.arange(max(0, N - 20), N),
             np.linspace(0, N - 1, count // 2).astype(np.int64),
             rng.integers(0, N, count // 2)]
    if L_bounds is not None:
        for b in L_bounds:
            parts.append(np.clip(np.arange(b - 2, b + 2), 0, N - 1))
    return np.unique(np.concatenate(parts)).astype(np.int64)

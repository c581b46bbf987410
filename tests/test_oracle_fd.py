"""Pins of the frequency-domain oracle (oracle/few_oracle_fd.c, the CPU branch
of get_waveform_generic_fd, src/few/cutils/interpolate.cu:1142-1563) and of
the kernel's per-bin code against it (host twin).

* SPAFunc against scipy's modified Bessel function: the routine equals
  -i sqrt(x) e^{-ix} K_1/3(-ix) (principal branches) up to the truncation of
  the reference's own series, measured per range of x;
* cube_roots against numpy.roots;
* the whole sum, as test_fd.py checks the reference (tests/test_fd.py:70-140):
  the overlap of h+ with the FFT of the time-domain waveform's real part, and
  its absolute scale: |h+|^2 = dt^2 |FFT|^2 band by band;
* the kernel's item/bin routines, run on the host in the oracle's loop order,
  equal the oracle bit for bit.
"""

import ctypes

import numpy as np
import pytest
import scipy.special as sp

import oracle
from fastemriwaveforms_amd import synthetic
from helpers import fd_frequencies, oracle_fd, oracle_fd_inputs, oracle_waveform

P = ctypes.c_void_p


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


# (x range, bound on |SPA - exact| / |exact|): the reference's series is
# truncated at x^26 (|x| <= 7) and its asymptotic series at x^-8
SPA_BOUNDS = [((1e-3, 1.0), 1e-14), ((1.0, 4.0), 1e-12), ((4.0, 6.0), 5e-8), ((6.0, 7.0), 3e-6),
              ((7.0 + 1e-9, 8.0), 1e-6), ((8.0, 12.0), 3e-7), ((12.0, 50.0), 1e-8),
              ((50.0, 1e4), 1e-13)]


@pytest.mark.parametrize("rng_x,bound", SPA_BOUNDS)
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_spa_func_vs_bessel(rng_x, bound, sign):
    x = sign * np.linspace(*rng_x, 400)
    spa = oracle.fd_spa_func(x)
    exact = -1j * np.sqrt(x.astype(complex)) * np.exp(-1j * x) * sp.kv(1.0 / 3.0, -1j * x)
    err = np.abs(spa - exact) / np.abs(exact)
    assert err.max() < bound, (rng_x, sign, err.max())


def test_cube_roots_vs_numpy():
    rng = np.random.default_rng(1)
    rows = [(rng.normal() * 1e-24, rng.normal() * 1e-17, abs(rng.normal()) * 1e-10 + 1e-11,
             rng.uniform(-3e-5, 3e-5)) for _ in range(300)]
    rows += [(1.0, -6.0, 11.0, -6.0), (2.0, -3.0, -11.0, 6.0), (1.0, 1.0, 1.0, 1.0)]
    rows = np.array(rows)
    got = oracle.fd_cube_roots(rows)
    for g, (a, b, c, d) in zip(got, rows):
        ref = np.roots([a, b, c, d])
        scale = np.abs(ref).max()
        assert max(np.abs(ref - r).min() for r in g) / scale < 1e-12


def test_cube_roots_degenerate_nan():
    """b = c = 0 makes Cardano's C vanish and 0/0 appear: the reference's
    formula yields NaN roots there (no bin is filled); the oracle keeps that."""
    r = oracle.fd_cube_roots([[1.0, 0.0, 0.0, -8.0]])
    assert np.isnan(r.real).all()


def _small(T=0.05, L=30, K=12, seed=5, chirp="smooth"):
    return synthetic.make_input(T, 10.0, L, "m_nonneg_30", K, seed, name="fd", chirp=chirp)


def test_fd_overlap_with_fft_of_td():
    """tests/test_fd.py:70-140 restated on synthetic inputs: the stationary-phase
    h+ overlaps the FFT of the time-domain real part to better than 1e-2."""
    inp = _small(T=0.05, L=40, K=16, seed=11, chirp="rising")
    h_td = oracle_waveform(inp)
    npts = inp.num_pts + (1 - inp.num_pts % 2)         # odd length, as the FD output
    h_td = np.concatenate([h_td, np.zeros(npts - len(h_td))])
    freq = oracle.fd_frequency_grid(len(h_td), 0, inp.dt)
    wave, _ = oracle_fd(inp, freq)
    hp = oracle.fd_post(wave, freq)[0]
    mask = freq >= 0.0
    ts = np.fft.fftshift(np.fft.fft(np.real(h_td)))[mask]
    fd = hp[mask]
    ac = np.vdot(ts, fd) / np.sqrt(np.vdot(ts, ts) * np.vdot(fd, fd))
    assert 1 - ac.real < 1e-2, 1 - ac.real


def test_fd_absolute_scale_vs_fft_of_td():
    """The stationary-phase h+ is the continuous Fourier transform of the
    time-domain real part: |h+(f)|^2 = dt^2 |FFT(Re h)(f)|^2 band by band
    (20 bands; those carrying 5 % of the power or more, which hold most of it,
    agree to 1 %; the total to 2 %).  The overlap check above cannot see a
    constant factor in the amplitude; this one does (a factor 2, sqrt(2) or
    2 pi moves the ratio far outside the bounds)."""
    inp = _small(T=0.05, L=40, K=16, seed=11, chirp="rising")
    h_td = oracle_waveform(inp)
    npts = inp.num_pts + (1 - inp.num_pts % 2)
    h_td = np.concatenate([h_td, np.zeros(npts - len(h_td))])
    freq = oracle.fd_frequency_grid(len(h_td), 0, inp.dt)
    wave, _ = oracle_fd(inp, freq)
    hp = oracle.fd_post(wave, freq)[0]
    mask = freq >= 0.0
    p_td = np.abs(np.fft.fftshift(np.fft.fft(np.real(h_td)))[mask]) ** 2 * inp.dt ** 2
    p_fd = np.abs(hp[mask]) ** 2
    assert p_fd.sum() / p_td.sum() == pytest.approx(1.0, abs=0.02)
    edges = np.linspace(0, len(p_td), 21).astype(int)
    covered = 0.0
    for a, b in zip(edges[:-1], edges[1:]):
        share = p_td[a:b].sum() / p_td.sum()
        if share >= 0.05:
            covered += share
            assert p_fd[a:b].sum() / p_td[a:b].sum() == pytest.approx(1.0, abs=0.01), (a, b)
    assert covered > 0.8


def test_fd_segment_inds_tile_the_band():
    """Monotonic in-band mode frequencies: each segment's index range starts one
    past its lower edge bin and the segments tile the mode's band without gaps."""
    inp = _small(chirp="rising")
    f_phi, f_theta, f_r = fd_frequencies(inp)
    freq = oracle.fd_frequency_grid(inp.num_pts, 1 - inp.num_pts % 2, inp.dt)
    start, end, zero = oracle.fd_segment_inds(inp.t, f_phi, f_theta, f_r, inp.m, inp.k, inp.n,
                                              freq)
    assert start.shape == (inp.K, inp.L - 1) and end.shape == start.shape
    assert zero == len(freq) // 2 and freq[zero] == 0.0
    df = freq[1] - freq[0]
    fmax = np.abs(inp.m[:, None] * f_phi + inp.n[:, None] * f_r).max(axis=1)
    for q in np.where(fmax < 0.9 * freq[-1])[0]:
        seg_f = inp.m[q] * f_phi + inp.k[q] * f_theta + inp.n[q] * f_r
        lo = np.minimum(seg_f[:-1], seg_f[1:])
        # start = trunc((f_lo - f_0)/df) + 1: the first bin at or above f_lo
        assert np.all(freq[start[q] - 1] <= lo + 1e-12) and np.all(freq[start[q]] >= lo - df)
        o = np.argsort(start[q])
        assert np.all(start[q][o][1:] == end[q][o][:-1] + 1)


@pytest.fixture(scope="module")
def tw(twin):
    twin.twin_fd_waveform.argtypes = [P, P, P, P, P, P, P, ctypes.c_int, P, ctypes.c_int,
                                      ctypes.c_int, P, P, P, ctypes.c_int, P, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int]
    twin.twin_fd_spa.argtypes = [P, ctypes.c_long, P]
    return twin


@pytest.mark.parametrize("include_minus_m,separate_modes", [(True, False), (False, False),
                                                            (True, True)])
def test_twin_fd_bitwise(tw, include_minus_m, separate_modes):
    """The kernel's fd_load_item / fd_bin, looped in the CPU branch's order,
    equal the oracle bit for bit (host libm on both sides)."""
    inp = _small(T=0.02, L=25, K=10, seed=3)
    freq = oracle.fd_frequency_grid(inp.num_pts, 1 - inp.num_pts % 2, inp.dt)
    ref, (start, end, zero) = oracle_fd(inp, freq, include_minus_m=include_minus_m,
                                        separate_modes=separate_modes)
    interp, pcoef = oracle_fd_inputs(inp)
    Nf = len(freq)
    wave = np.zeros((inp.K, Nf) if separate_modes else (Nf,), dtype=np.complex128)
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    m, k, n, s0, s1 = i32(inp.m), i32(inp.k), i32(inp.n), i32(start), i32(end)
    ylm = np.ascontiguousarray(inp.ylms)
    rc = tw.twin_fd_waveform(_p(wave), _p(interp), _p(inp.t), _p(pcoef), _p(m), _p(k), _p(n),
                             inp.K, _p(inp.t), inp.L, Nf, _p(freq), _p(s0), _p(s1), inp.L - 1,
                             _p(ylm), zero, int(include_minus_m), int(separate_modes))
    assert rc == 0
    assert np.abs(ref).max() > 0
    assert np.array_equal(wave, ref)


def test_twin_spa_bitwise(tw):
    x = np.concatenate([np.linspace(-60, 60, 2001), [7.0, -7.0, 7.0000001]])
    out = np.zeros(len(x), dtype=np.complex128)
    tw.twin_fd_spa(_p(x), len(x), _p(out))
    assert np.array_equal(out, oracle.fd_spa_func(x))

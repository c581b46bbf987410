"""GPU parity of the frequency-domain interpolated sum (csrc/fd.hip,
few_get_waveform_generic_fd*) and of FDInterpolatedModeSum's glue against the
oracle (oracle/few_oracle_fd.c, oracle.fd_segment_inds, oracle.fd_post).

Bars:
* segment-to-bin indices: identical integers;
* the stationary-phase sum: rel-L2 below FD_TOL against the oracle on the
  same inputs.  The operations are the oracle's (contraction off); what
  differs is the device libm (hypot, atan2, log, exp, sincos) against glibc in
  the last bit, which Cardano's cube root and the phase exponential carry
  into the result.  The summed output is bin-owned (each bin written by one
  thread in a fixed order), so it is also identical from run to run.
"""

import numpy as np
import pytest

import oracle
from fastemriwaveforms_amd import synthetic
from helpers import fd_frequencies, oracle_fd, oracle_fd_inputs

pytestmark = pytest.mark.gpu

FD_TOL = 2e-11


def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _inp(T=0.05, L=40, K=16, seed=11, chirp="rising"):
    return synthetic.make_input(T, 10.0, L, "m_nonneg_30", K, seed, name="fd", chirp=chirp)


def _grid(inp):
    return oracle.fd_frequency_grid(inp.num_pts, 1 - inp.num_pts % 2, inp.dt)


def _run_capi(inp, freq, start, end, zero, include_minus_m=True, separate_modes=False,
              wave0=None):
    import ctypes
    import torch
    from fastemriwaveforms_amd._lib import check, lib
    dev = torch.device("cuda", torch.cuda.current_device())
    interp, pcoef = oracle_fd_inputs(inp)
    Nf = len(freq)
    d = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)
    shape = (inp.K, Nf) if separate_modes else (Nf,)
    wave = (torch.zeros(shape, dtype=torch.complex128, device=dev) if wave0 is None
            else d(wave0, torch.complex128))
    args = [d(interp), d(inp.t), d(pcoef), d(inp.m, torch.int32), d(inp.k, torch.int32),
            d(inp.n, torch.int32), d(inp.t), d(freq), d(start, torch.int32), d(end, torch.int32),
            d(inp.ylms, torch.complex128)]
    nbytes = ctypes.c_size_t(0)
    check(lib().few_fd_workspace_bytes(inp.K, inp.L - 1, ctypes.byref(nbytes)), "ws")
    ws = torch.empty(max(1, nbytes.value), dtype=torch.uint8, device=dev)
    a = [x.data_ptr() for x in args]
    rc = lib().few_get_waveform_generic_fd_async(
        wave.data_ptr(), a[0], a[1], a[2], a[3], a[4], a[5], inp.K, inp.dt, a[6], inp.L, Nf, a[7],
        a[8], a[9], inp.L - 1, a[10], zero, int(include_minus_m), int(separate_modes),
        ws.data_ptr(), ws.numel(), int(torch.cuda.current_stream().cuda_stream))
    check(rc, "fd sum")
    torch.cuda.synchronize()
    return wave.cpu().numpy()


@pytest.mark.parametrize("case", ["rising", "smooth", "rising_long"])
@pytest.mark.parametrize("include_minus_m,separate_modes", [(True, False), (False, False),
                                                            (False, True)])
def test_fd_kernel_vs_oracle(gpu, case, include_minus_m, separate_modes):
    inp = {"rising": lambda: _inp(),
           "smooth": lambda: _inp(chirp="smooth", seed=5),
           "rising_long": lambda: _inp(T=0.3, L=80, K=24, seed=21)}[case]()
    freq = _grid(inp)
    ref, (start, end, zero) = oracle_fd(inp, freq, include_minus_m=include_minus_m,
                                        separate_modes=separate_modes)
    got = _run_capi(inp, freq, start, end, zero, include_minus_m, separate_modes)
    err = _rel_l2(got, ref)
    print(f"[fd] {case} minus={include_minus_m} separate={separate_modes}: rel-L2 {err:.3e} "
          f"({int((np.abs(ref) > 0).sum())} nonzero bins)")
    assert np.abs(ref).max() > 0
    assert err < FD_TOL


def test_fd_kernel_accumulates(gpu):
    """The waveform is accumulated into (interpolate.cu:1478-1490)."""
    inp = _inp(T=0.02, L=25, K=8, seed=3)
    freq = _grid(inp)
    ref, (start, end, zero) = oracle_fd(inp, freq)
    rng = np.random.default_rng(0)
    w0 = rng.normal(size=len(freq)) + 1j * rng.normal(size=len(freq))
    got = _run_capi(inp, freq, start, end, zero, wave0=w0)
    assert _rel_l2(got - w0, ref) < FD_TOL


def test_fd_kernel_empty_and_out_of_range(gpu):
    """Empty ranges (start > end) add nothing; ranges past either end of the
    array are clipped to it."""
    inp = _inp(T=0.02, L=25, K=8, seed=3)
    freq = _grid(inp)
    _, (start, end, zero) = oracle_fd(inp, freq)
    s2, e2 = start.copy(), end.copy()
    s2[0, :] = 5
    e2[0, :] = 4                    # empty
    s2[1, 0], e2[1, 0] = -50, 3     # starts before bin 0
    s2[2, 1], e2[2, 1] = len(freq) - 3, len(freq) + 40  # runs past the end
    f_phi, f_theta, f_r = fd_frequencies(inp)
    interp, pcoef = oracle_fd_inputs(inp)
    ref = oracle.fd_waveform(interp, inp.t, pcoef, inp.m, inp.k, inp.n, inp.t, freq, s2, e2,
                             inp.ylms, zero)
    got = _run_capi(inp, freq, s2, e2, zero)
    assert _rel_l2(got, ref) < FD_TOL


def test_fd_bin_owned_deterministic(gpu):
    """Two runs of the summed output are identical bit for bit (no atomics)."""
    inp = _inp(T=0.3, L=80, K=24, seed=21)
    freq = _grid(inp)
    ref, (start, end, zero) = oracle_fd(inp, freq)
    a = _run_capi(inp, freq, start, end, zero)
    b = _run_capi(inp, freq, start, end, zero)
    assert np.array_equal(a, b)
    assert _rel_l2(a, ref) < FD_TOL


@pytest.mark.parametrize("zero_shift", [-7, 0, 5, "first", "last"])
def test_fd_bin_owned_zero_index_and_grid(gpu, zero_shift):
    """Bin ownership pairs bin b with its mirror 2z - b: grids and zero indices
    where some bins have no partner (zero index off-centre, at the first or the
    last bin) give the oracle's sum, whose -f terms land at 2z - i when inside."""
    inp = _inp(T=0.02, L=25, K=8, seed=3)
    freq = _grid(inp)
    _, (start, end, zero) = oracle_fd(inp, freq)
    z = {"first": 0, "last": len(freq) - 1}.get(zero_shift, None)
    if z is None:
        z = zero + zero_shift
    f_phi, f_theta, f_r = fd_frequencies(inp)
    interp, pcoef = oracle_fd_inputs(inp)
    ref = oracle.fd_waveform(interp, inp.t, pcoef, inp.m, inp.k, inp.n, inp.t, freq, start, end,
                             inp.ylms, z)
    got = _run_capi(inp, freq, start, end, z)
    assert np.abs(ref).max() > 0
    assert _rel_l2(got, ref) < FD_TOL, zero_shift


def _turnover_freqs(inp, bump):
    """f_r with a bump that gives the mode frequencies m f_phi + n f_r of some
    modes an interior extremum (a turnover)."""
    f_phi, f_theta, f_r = fd_frequencies(inp)
    tau = inp.t / inp.t[-1]
    return f_phi, f_theta, f_r + bump * np.exp(-((tau - 0.55) / 0.12) ** 2)


@pytest.mark.parametrize("chirp,bump,modes", [("rising", 0.0, None), ("smooth", 0.0, None),
                                               ("rising", 3e-4, "one"), ("rising", 3e-4, None)])
def test_segment_inds_match_oracle(gpu, chirp, bump, modes):
    """FDInterpolatedModeSum's index bookkeeping (torch on the device) gives the
    oracle's integers, including the turnover correction, which the reference
    applies only when exactly one turnover exists in total."""
    import torch
    from fastemriwaveforms_amd.fdinterp import segment_frequency_inds
    inp = _inp(chirp=chirp, K=40, seed=13)
    f = _turnover_freqs(inp, bump)
    m, k, n = inp.m, inp.k, inp.n
    if modes == "one":
        # keep the first mode that turns over on the reference's 5000-point grid
        # (fdinterp.py:241-262) plus modes that do not
        fs, _, _ = oracle.interpolate_arrays(inp.t, np.asarray(f))
        t_new = np.linspace(inp.t.min(), inp.t.max(), 5000)
        nf = oracle.spline_eval(np.tile(inp.t, (3, 1)), fs, t_new)
        pm = m[:, None] * nf[0] + k[:, None] * nf[1] + n[:, None] * nf[2]
        tx, tn = t_new[pm.argmax(-1)], t_new[pm.argmin(-1)]
        turns = ((tx > 0) & (tx < inp.t[-1])) | ((tn > 0) & (tn < inp.t[-1]))
        assert turns.sum() >= 2
        keep = [int(np.argmax(turns))] + [q for q in range(len(m)) if not turns[q]]
        m, k, n = m[keep], k[keep], n[keep]
    freq = _grid(inp)
    s_ref, e_ref, z_ref = oracle.fd_segment_inds(inp.t, *f, m, k, n, freq)
    s, e, z = segment_frequency_inds(inp.t, *f, m, k, n,
                                     torch.from_numpy(freq).to("cuda"))
    assert z == z_ref
    assert np.array_equal(s.cpu().numpy(), s_ref)
    assert np.array_equal(e.cpu().numpy(), e_ref)


def test_fd_interpolated_mode_sum_end_to_end(gpu):
    """FDInterpolatedModeSum(...)(t, ...) against the oracle pipeline:
    splines, indices, the sum and the h+/hx assembly (fdinterp.py:167-542)."""
    import torch
    from fastemriwaveforms_amd.fdinterp import FDInterpolatedModeSum
    inp = _inp(T=0.05, L=40, K=16, seed=11)
    f = fd_frequencies(inp)
    pc, _, _ = oracle.interpolate_arrays(inp.t, inp.phase_knots)
    dense = oracle.cubic_to_dense(pc, inp.t)                     # (8, L-1, 3)
    phase_coeffs = np.ascontiguousarray(np.transpose(dense, (1, 2, 0)))  # (L-1, 3, 8)
    ylms_in = inp.ylms.copy()
    ylms_in[inp.K:] *= (-1.0) ** inp.l                           # the class applies (-1)^l
    summ = FDInterpolatedModeSum(output_type="fd", pad_output=True)
    out = summ(inp.t, inp.teuk, ylms_in, inp.t, phase_coeffs, inp.l, inp.m, inp.k, inp.n,
               T=inp.T_years, dt=inp.dt, frequencies=f)
    got = out.cpu().numpy()
    freq = summ.frequency.cpu().numpy()
    ref_two_sided, _ = oracle_fd(inp, freq)
    ref = oracle.fd_post(ref_two_sided, freq)
    assert got.shape == ref.shape == (2, len(freq))
    err = _rel_l2(got, ref)
    print(f"[fd] end-to-end rel-L2 {err:.3e}, Nf={len(freq)}")
    assert err < FD_TOL
    out_pos = FDInterpolatedModeSum(output_type="fd")(
        inp.t, inp.teuk, ylms_in, inp.t, phase_coeffs, inp.l, inp.m, inp.k, inp.n,
        T=inp.T_years, dt=inp.dt, frequencies=f, mask_positive=True)
    assert tuple(out_pos.shape) == (2, int((freq >= 0).sum()))
    assert _rel_l2(out_pos.cpu().numpy(), ref[:, freq >= 0]) < FD_TOL


def test_fd_dropin_wrapper_numpy(gpu):
    """get_waveform_generic_fd_wrap with the reference's positional arguments on
    numpy arrays (the numpy-featured drop-in), written back in place."""
    from fastemriwaveforms_amd import pyinterp
    inp = _inp(T=0.02, L=25, K=8, seed=3)
    freq = _grid(inp)
    ref, (start, end, zero) = oracle_fd(inp, freq)
    interp, pcoef = oracle_fd_inputs(inp)
    wave = np.zeros(len(freq), dtype=np.complex128)
    pyinterp.get_waveform_generic_fd_wrap(
        wave, interp.flatten(), inp.t.copy(), pcoef.flatten(), inp.m.astype(np.int32),
        inp.k.astype(np.int32), inp.n.astype(np.int32), inp.K, inp.dt, inp.t.copy(), inp.L,
        len(freq), freq, start.astype(np.int32).flatten(), end.astype(np.int32).flatten(),
        inp.L - 1, inp.ylms.copy(), zero, True, False)
    assert _rel_l2(wave, ref) < FD_TOL

"""GPU parity of the AAK summation (csrc/aak.hip, few_aak_waveform_async /
few_aak_get_waveform, AAKSummation) against the oracle
(oracle/few_oracle_aak.c).  The operations are the oracle's (contraction off,
the same Numerical Recipes Bessel routines, sincos for every cos/sin pair);
the device libm (sincos, acos, atan2, pow, sqrt) differs from glibc in the
last bit: rel-L2 < 1e-12."""

import numpy as np
import pytest

from helpers import aak_input, oracle_aak

pytestmark = pytest.mark.gpu

AAK_TOL = 1e-12


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("mich", [False, True])
@pytest.mark.parametrize("case", ["small", "eccentric", "circular", "no_spin"])
def test_aak_summation_vs_oracle(gpu, case, mich):
    from fastemriwaveforms_amd.aakwave import AAKSummation
    kw = {"small": dict(T_years=0.01, L=15, seed=1),
          "eccentric": dict(T_years=0.05, L=40, seed=2, e0=0.7, e1=0.5),
          "circular": dict(T_years=0.02, L=20, seed=3, e0=0.0, e1=0.0),
          "no_spin": dict(T_years=0.02, L=20, seed=4)}[case]
    t, c, prm, n_out = aak_input(**kw)
    if case == "no_spin":
        prm = dict(prm, S=0.0)
    nmodes = 25 if case == "eccentric" else 12
    ref = oracle_aak(t, c, prm, n_out, nmodes=nmodes, mich=mich)
    summ = AAKSummation()
    out = summ(t, prm["M"], prm["S"], prm["dist"], prm["mu"], prm["qS"], prm["phiS"], prm["qK"],
               prm["phiK"], nmodes, t, c, mich=mich, T=kw["T_years"], dt=10.0)
    got = out.cpu().numpy()[:n_out]
    err = _rel(got, ref)
    print(f"[aak] {case} mich={mich}: N={n_out} rel-L2 {err:.3e}")
    assert np.abs(ref).max() > 0
    assert err < AAK_TOL


def test_aak_dropin_numpy(gpu):
    """pyWaveform with the reference's positional arguments on numpy arrays
    (gpuAAKWrap.pyx:17-33), h_t on the host."""
    from fastemriwaveforms_amd import pyinterp
    t, c, prm, n_out = aak_input(T_years=0.01, L=15, seed=7)
    ref = oracle_aak(t, c, prm, n_out, nmodes=10)
    wave = np.zeros(n_out, dtype=np.complex128)
    dense = np.ascontiguousarray(np.transpose(c, (2, 0, 1))).flatten()
    pyinterp.pyWaveform(wave, dense, prm["M"], prm["S"], prm["mu"], prm["qS"], prm["phiS"],
                        prm["qK"], prm["phiK"], prm["dist"], 10, False, len(t), n_out, 10.0,
                        t.copy())
    assert _rel(wave, ref) < AAK_TOL


def test_aak_samples_past_the_trajectory_untouched(gpu):
    """Samples past the last used segment keep their values (the reference
    overwrites only the segments' samples)."""
    import torch
    from fastemriwaveforms_amd._lib import check, lib
    t, c, prm, n_out = aak_input(T_years=0.01, L=15, seed=8)
    dev = torch.device("cuda", torch.cuda.current_device())
    extra = 500
    wave = torch.full((n_out + extra,), 3.0 + 1.0j, dtype=torch.complex128, device=dev)
    dense = torch.from_numpy(np.ascontiguousarray(np.transpose(c, (2, 0, 1)))).to(dev)
    h_t = torch.from_numpy(t).to(dev)
    ws = torch.empty(int(lib().few_aak_workspace_bytes(len(t))), dtype=torch.uint8, device=dev)
    check(lib().few_aak_waveform_async(wave.data_ptr(), dense.data_ptr(), prm["M"], prm["S"],
                                       prm["mu"], prm["qS"], prm["phiS"], prm["qK"], prm["phiK"],
                                       prm["dist"], 6, 0, len(t), n_out, 10.0, h_t.data_ptr(),
                                       ws.data_ptr(), ws.numel(),
                                       int(torch.cuda.current_stream().cuda_stream)), "aak")
    got = wave.cpu().numpy()
    ref = oracle_aak(t, c, prm, n_out, nmodes=6)
    assert _rel(got[:n_out], ref) < AAK_TOL
    assert np.all(got[n_out:] == 3.0 + 1.0j)

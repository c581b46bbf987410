"""GPU parity on the trajectory's own phase format: DOPR853-style dense output
with non-zero rcont5..8 on all three phases (Phi_theta != 0; the trajectory
hands these over, trajectory/inspiral.py:143-147, and get_waveform evaluates
all eight terms, interpolate.cu:637-644), through every sum that reads them:
the FP64 direct sum, the mixed FP32 and split-FP16 factorised sums (with and
without k != 0 modes), the batch, the drop-in InterpolatedModeSum, the
adjoint's amplitude gradients, the frequency-domain sum (whose f-dot and
f-ddot come from the same polynomial) and the AAK sum.

Bars: FP64 rel-L2 <= 1e-12 against the oracle on the same coefficients; mixed
paths < 1e-6; FD 2e-11 (device libm in the last bit); AAK 1e-12.
"""

import numpy as np
import pytest
import torch

import oracle
from fastemriwaveforms_amd import synthetic
from helpers import (aak_input, oracle_aak, oracle_coefficients, oracle_fd, oracle_fd_inputs,
                     oracle_waveform, spread_indices)

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-6


def _inp(modes="m_nonneg_30", K=300, T=0.02, L=40, seed=41):
    inp = synthetic.make_input(T, 10.0, L, modes, K, seed, phases="dopr")
    assert inp.dense is not None and np.all(np.abs(inp.dense[4:]) > 0)
    assert np.all(inp.phase_knots[1, 1:] != 0)
    return inp


def _tol(prec):
    return FP64_TOL if prec == "fp64" else MIXED_TOL


@pytest.mark.parametrize("modes", ["m_nonneg_30", "k1_20"])
@pytest.mark.parametrize("prec", ["fp64", "mixed", "mixed16"])
def test_dopr_mode_sum(gpu, modes, prec):
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    inp = _inp(modes, K=400 if modes == "k1_20" else 300)
    if modes == "k1_20":
        assert np.any(inp.k != 0)
    p = ModeSumProblem(inp, gpu)
    if prec != "fp64":
        assert p.plan.supported and (prec == "mixed" or p.plan.supported16)
    h = p.run(prec).cpu().numpy()
    ref = oracle_waveform(inp)
    err = oracle.rel_l2(h, ref)
    print(f"dopr {modes} {prec}: N={inp.num_pts} K={inp.K} rel-L2 {err:.3e}")
    assert err <= _tol(prec)


@pytest.mark.parametrize("prec", ["fp64", "mixed16"])
def test_dopr_cpu0_every_sample(gpu, prec):
    """The cpu0 shape (0.1 yr, 100 knots, 500 modes) with DOPR phases: every sample."""
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    inp = synthetic.make_input(0.1, 10.0, 100, "m_nonneg_30", 500, 1000, phases="dopr")
    h = ModeSumProblem(inp, gpu).run(prec).cpu().numpy()
    err = oracle.rel_l2(h, oracle_waveform(inp))
    print(f"dopr cpu0 {prec}: rel-L2 {err:.3e}")
    assert err <= _tol(prec)


def test_dopr_full1_mixed16(gpu):
    """The full1 shape (1 yr, 120 knots, 3843 modes) with DOPR phases: mixed16
    against the FP64 GPU path on every sample and the FP64 path against the
    oracle on spread samples."""
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    inp = synthetic.make_input(1.0, 10.0, 120, "m_nonneg_30", None, 1001, phases="dopr")
    p = ModeSumProblem(inp, gpu)
    h64 = p.run("fp64").cpu().numpy()
    h16 = p.run("mixed16").cpu().numpy()
    idx = spread_indices(inp.num_pts, count=1200)
    ref = oracle_waveform(inp, idx)
    e64 = oracle.rel_l2(h64[idx], ref)
    e16 = oracle.rel_l2(h16, h64)
    print(f"dopr full1: fp64 vs oracle {e64:.3e}, mixed16 vs fp64 {e16:.3e}")
    assert e64 <= FP64_TOL and e16 < MIXED_TOL


@pytest.mark.parametrize("prec", ["fp64", "mixed16"])
def test_dopr_dropin_class(gpu, prec):
    """InterpolatedModeSum.__call__ with the trajectory's (L-1, 3, 8) coefficients."""
    from fastemriwaveforms_amd.summation import InterpolatedModeSum
    inp = _inp(K=200, seed=43)
    dopr = np.transpose(inp.dense, (1, 2, 0))
    ylms = inp.ylms.copy()
    ylms[inp.K:] *= (-1.0) ** inp.l  # sum() applies (-1)^l again (interpolatedmodesum.py:356-358)
    summ = InterpolatedModeSum(precision=prec)
    h = summ(inp.t, inp.teuk, ylms, inp.t, dopr, inp.l, inp.m, inp.k, inp.n,
             T=inp.T_years, dt=inp.dt).cpu().numpy()
    err = oracle.rel_l2(h, oracle_waveform(inp))
    print(f"dopr drop-in {prec}: rel-L2 {err:.3e}")
    assert err <= _tol(prec)


@pytest.mark.parametrize("prec", ["fp64", "mixed16"])
def test_dopr_batch(gpu, prec):
    """Uneven batch with dense-output phases in pcoef (phase_knots NULL): every
    waveform against the oracle."""
    from fastemriwaveforms_amd.pipeline import BatchProblem
    rng = np.random.default_rng(77)
    inputs = [synthetic.make_input(float(rng.uniform(0.002, 0.01)), 10.0,
                                   int(rng.integers(8, 30)), "m_nonneg_30",
                                   int(rng.integers(20, 200)), 500 + b, phases="dopr")
              for b in range(5)]
    bp = BatchProblem(inputs, gpu)
    assert bp.dense_input
    bp.run(prec)
    for b, x in enumerate(inputs):
        err = oracle.rel_l2(bp.waveform(b).cpu().numpy(), oracle_waveform(x))
        print(f"dopr batch {prec} waveform {b}: rel-L2 {err:.3e}")
        assert err <= _tol(prec)


def test_dopr_adjoint_amplitudes_fp64(gpu):
    """The FP64 adjoint recomputes the phases from all eight coefficients: its
    amplitude-coefficient gradients match the oracle's adjoint."""
    from fastemriwaveforms_amd import _lib
    from fastemriwaveforms_amd._lib import check, lib
    inp = _inp(K=60, T=0.004, L=16, seed=45)
    interp_tl, pcoef = oracle_coefficients(inp)
    rng = np.random.default_rng(4)
    g = rng.normal(size=inp.num_pts) + 1j * rng.normal(size=inp.num_pts)
    gi_ref, _ = oracle.waveform_adjoint(g, interp_tl, inp.t, pcoef, inp.m, inp.k, inp.n, inp.L,
                                        inp.num_pts, inp.K, inp.ylms, inp.dt, inp.t)
    d = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(gpu, dt)  # noqa
    L, K, N = inp.L, inp.K, inp.num_pts
    gi = torch.zeros((4, L, 2 * K), dtype=torch.float64, device=gpu)
    gp = torch.zeros((4, L, 3), dtype=torch.float64, device=gpu)
    nb = int(lib().few_adjoint_workspace_bytes(L, N, K, _lib.FEW_PRECISION_FP64))
    ws = torch.empty(nb, dtype=torch.uint8, device=gpu)
    args = [d(g, torch.complex128), d(interp_tl), d(inp.t), d(pcoef), d(inp.m, torch.int32),
            d(inp.k, torch.int32), d(inp.n, torch.int32), d(inp.ylms, torch.complex128)]
    rc = lib().few_waveform_adjoint_async(
        args[0].data_ptr(), args[1].data_ptr(), args[2].data_ptr(), args[3].data_ptr(),
        args[4].data_ptr(), args[5].data_ptr(), args[6].data_ptr(), L, N, K, args[7].data_ptr(),
        inp.dt, args[2].data_ptr(), gi.data_ptr(), gp.data_ptr(), _lib.FEW_PRECISION_FP64,
        ws.data_ptr(), ws.numel(), int(torch.cuda.current_stream().cuda_stream))
    check(rc, "waveform_adjoint")
    err = oracle.rel_l2(gi.cpu().numpy(), np.asarray(gi_ref).reshape(4, L, 2 * K))
    print(f"dopr adjoint amplitudes: rel-L2 {err:.3e}")
    assert err <= FP64_TOL


def test_dopr_frequency_domain(gpu):
    """The FD sum reads the phase polynomial and its first two derivatives
    (f-dot, f-ddot) from the dense output: DOPR phases against the oracle."""
    from fastemriwaveforms_amd.fdinterp import FDInterpolatedModeSum
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    inp = synthetic.make_input(0.05, 10.0, 40, "m_nonneg_30", 16, 11, phases="dopr")
    pc = ModeSumProblem(inp, gpu)
    ylms_in = inp.ylms.copy()
    ylms_in[inp.K:] *= (-1.0) ** inp.l
    summ = FDInterpolatedModeSum(output_type="fd")
    hf = summ(inp.t, inp.teuk, ylms_in, inp.t, pc.pcoef.permute(1, 2, 0).contiguous(), inp.l,
              inp.m, inp.k, inp.n, T=inp.T_years, dt=inp.dt, frequencies=tuple(inp.freqs))
    freq = summ.frequency.cpu().numpy()
    ref = oracle.fd_post(oracle_fd(inp, freq)[0], freq)
    err = oracle.rel_l2(hf.cpu().numpy(), ref)
    print(f"dopr FD: Nf={len(freq)} rel-L2 {err:.3e}")
    assert np.abs(ref).max() > 0 and err < 2e-11


@pytest.mark.parametrize("mich", [False, True])
def test_dopr_aak(gpu, mich):
    """AAK with DOPR853-style dense output of p, e, Y and the three phases."""
    from fastemriwaveforms_amd.aakwave import AAKSummation
    t, c, prm, n_out = aak_input(T_years=0.03, L=30, seed=21, dopr=True)
    assert np.all(np.abs(c[:, :, 4:]) > 0)
    ref = oracle_aak(t, c, prm, n_out, nmodes=15, mich=mich)
    summ = AAKSummation()
    out = summ(t, prm["M"], prm["S"], prm["dist"], prm["mu"], prm["qS"], prm["phiS"], prm["qK"],
               prm["phiK"], 15, t, c, mich=mich, T=0.03, dt=10.0)
    got = out.cpu().numpy()[:n_out]
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print(f"dopr AAK mich={mich}: rel-L2 {err:.3e}")
    assert np.abs(ref).max() > 0 and err < 1e-12

"""GPU tests of the FP64 factorised sum (FEW_PRECISION_FACT64): the tiles,
phasors, MFMA products and accumulation are all FP64, so it differs from the
reference arithmetic only by the rounding of the phase products (m Phi_phi is
rounded once in the reference, and as m steps of a phasor recurrence here):
relative ~1e-10 at phases of 1e5-1e7 rad, against ~1.5e-7 for the split-FP16
sum.  Bar FACT64_TOL below (measured values printed)."""

import numpy as np
import pytest
import torch

import oracle
from helpers import oracle_waveform, spread_indices

pytestmark = pytest.mark.gpu

FACT64_TOL = 1e-8


def _run(inp, variant=0, stage=3, overwrite=False, wave0=None):
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    p = ModeSumProblem(inp)
    assert p.plan.supported64
    p.variant = variant
    p.build_splines()
    if wave0 is None:
        p.wave.zero_()
    else:
        p.wave.copy_(wave0)
    p.mode_sum("fact64", overwrite=overwrite)
    return p.wave.cpu().numpy(), p


@pytest.mark.parametrize("name", ["cpu0", "dopr_k", "n55"])
def test_fact64_vs_oracle(gpu, name):
    from fastemriwaveforms_amd import synthetic
    inp = {"cpu0": lambda: synthetic.make_workload("cpu0"),
           "dopr_k": lambda: synthetic.make_input(0.004, 10.0, 16, "k1_20", 200, 47, phases="dopr",
                                                  name="dopr_k"),
           "n55": lambda: synthetic.make_input(0.004, 10.0, 20, "m_all_55", 400, 48,
                                               name="n55")}[name]()
    h, _ = _run(inp)
    ref = oracle_waveform(inp)
    err = oracle.rel_l2(h, ref)
    print(f"[fact64] {name}: rel-L2 vs oracle {err:.3e}")
    assert err < FACT64_TOL


def test_fact64_full1_vs_fp64(gpu):
    """full1 (3843 modes, 3.16e6 samples): every sample against the FP64 direct
    path, and spread samples against the oracle."""
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    inp = synthetic.make_workload("full1")
    p = ModeSumProblem(inp)
    h64 = p.run("fp64").cpu().numpy()
    hf = p.run("fact64").cpu().numpy()
    err = oracle.rel_l2(hf, h64)
    idx = spread_indices(inp.num_pts, count=600)
    err_o = oracle.rel_l2(hf[idx], oracle_waveform(inp, idx))
    print(f"[fact64] full1: rel-L2 vs fp64 GPU {err:.3e}, vs oracle (600 samples) {err_o:.3e}")
    assert err < FACT64_TOL and err_o < FACT64_TOL


@pytest.mark.parametrize("tail", ["after_trajectory", "output_cut"])
def test_fact64_overwrite_and_variants(gpu, tail):
    """Overwrite (stage flag 4) gives what zeroing and accumulating gives, whatever
    the buffer held, also in the tail after the last knot or with the output cut
    short; variants 0 and 4 (16 / 32 samples per wave) give the same bits; the sum
    accumulates into a non-zero buffer."""
    from fastemriwaveforms_amd import synthetic
    inp = synthetic.small_input(seed=29, L=14, K=30, T_years=0.002)
    if tail == "after_trajectory":
        inp.num_pts += 3000
    else:
        inp.num_pts = int(0.5 * (inp.t[-4] + inp.t[-3]) / inp.dt)
    ref, p = _run(inp)
    gen = torch.Generator(device="cuda").manual_seed(5)
    junk = torch.view_as_complex(torch.randn(inp.num_pts, 2, dtype=torch.float64, device="cuda",
                                             generator=gen))
    got, _ = _run(inp, overwrite=True, wave0=junk)
    assert np.array_equal(got, ref)
    got4, _ = _run(inp, variant=4)
    assert np.array_equal(got4, ref)
    acc, _ = _run(inp, wave0=junk)
    assert np.allclose(acc - junk.cpu().numpy(), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    if tail == "after_trajectory":
        assert np.count_nonzero(got[-2000:]) == 0
    assert oracle.rel_l2(ref, oracle_waveform(inp)) < FACT64_TOL


def test_fact64_unknown_variant_refused(gpu):
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd._lib import FewRocmError
    with pytest.raises(FewRocmError):
        _run(synthetic.small_input(seed=3, L=10, K=12, T_years=0.001), variant=16)


def test_fact64_differentiable_pair(gpu):
    """The FP64 factorised forward with the FP64 factorised backward: gradients
    with respect to every amplitude and phase knot equal the FP64 pair's (the
    reference arithmetic forward, the per-mode FP64 adjoint) to FP64 level."""
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.autograd import DifferentiableModeSum
    inp = synthetic.small_input(seed=33, L=16, K=60, T_years=0.004)
    rng = np.random.default_rng(4)
    d = torch.from_numpy(rng.normal(size=inp.num_pts) + 1j * rng.normal(size=inp.num_pts)).cuda()
    grads = {}
    for prec, bwd in (("fp64", "fp64"), ("fact64", "fact64")):
        eng = DifferentiableModeSum(inp.t, inp.m, inp.k, inp.n, inp.ylms, inp.dt, inp.num_pts,
                                    precision=prec, backward=bwd)
        teuk = torch.from_numpy(np.ascontiguousarray(inp.teuk)).cuda().requires_grad_(True)
        pk = torch.from_numpy(np.ascontiguousarray(inp.phase_knots)).cuda().requires_grad_(True)
        h = eng(teuk, pk)
        loss = torch.real(torch.sum(torch.conj(d) * h))
        loss.backward()
        grads[prec] = (teuk.grad.cpu().numpy(), pk.grad.cpu().numpy(), h.detach().cpu().numpy())
    ea = oracle.rel_l2(grads["fact64"][0], grads["fp64"][0])
    ep = oracle.rel_l2(grads["fact64"][1], grads["fp64"][1])
    eh = oracle.rel_l2(grads["fact64"][2], grads["fp64"][2])
    print(f"[fact64 pair] forward {eh:.2e}, amplitude gradients {ea:.2e}, phase gradients {ep:.2e}")
    assert eh < FACT64_TOL and ea < 1e-9 and ep < 1e-6


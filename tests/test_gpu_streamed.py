"""GPU tests of the row-streamed variants of the split-FP16 sum (variants 16
and 20: B fragments held in registers, row blocks contracted one at a time
with the epilogue's Horner sum folded in between).  They perform the same
products, accumulations and roundings as variants 0 and 4, so their output is
compared bit for bit; cases the streamed kernel does not take (rows not in
arithmetic progression, more than 8 column steps) run the regular kernel."""

import numpy as np
import pytest
import torch

import oracle
from helpers import oracle_waveform

pytestmark = pytest.mark.gpu


def _single(inp, variant, stage=3, overwrite=False, junk=False):
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    p = ModeSumProblem(inp)
    p.variant = variant
    p.build_splines()
    if junk:
        gen = torch.Generator(device=p.wave.device).manual_seed(3)
        p.wave.copy_(torch.view_as_complex(torch.randn(inp.num_pts, 2, dtype=torch.float64,
                                                       device=p.wave.device, generator=gen)))
    else:
        p.wave.zero_()
    p.mode_sum("mixed16", overwrite=overwrite)
    return p.wave.cpu().numpy(), p


def _inputs():
    from fastemriwaveforms_amd import synthetic
    cut = synthetic.small_input(seed=29, L=14, K=30, T_years=0.002)
    cut.num_pts = int(0.5 * (cut.t[-4] + cut.t[-3]) / cut.dt)
    long = synthetic.small_input(seed=30, L=14, K=30, T_years=0.002)
    long.num_pts += 3000
    return {
        "cpu0": synthetic.make_workload("cpu0"),                                   # 2 column steps
        "n55": synthetic.make_input(0.003, 10.0, 20, "m_all_55", 400, 41, name="n55"),  # 4
        "k1_20": synthetic.make_input(0.003, 10.0, 20, "k1_20", 300, 42, name="k1_20"),
        "one_mode": synthetic.make_input(0.002, 10.0, 10, "m_nonneg_30", 1, 43, name="one"),
        "output_cut": cut,
        "after_trajectory": long,
    }


@pytest.mark.parametrize("name", ["cpu0", "n55", "k1_20", "one_mode", "output_cut",
                                  "after_trajectory"])
@pytest.mark.parametrize("streamed,regular", [(16, 0), (20, 4)])
def test_streamed_equals_regular(gpu, name, streamed, regular):
    inp = _inputs()[name]
    ref, p = _single(inp, regular)
    print(f"{name}: rows {p.plan.rows} cols {p.plan.cols}")
    got, _ = _single(inp, streamed)
    assert np.array_equal(got, ref), name
    # overwrite of a buffer holding junk gives the same samples
    got_ow, _ = _single(inp, streamed, overwrite=True, junk=True)
    assert np.array_equal(got_ow, ref), name
    if name == "cpu0":
        assert oracle.rel_l2(got, oracle_waveform(inp)) < 1e-6


@pytest.mark.parametrize("variant", [16, 20])
def test_streamed_batch_equals_regular(gpu, variant):
    from fastemriwaveforms_amd.pipeline import BatchProblem
    from test_gpu_batch import _small_batch
    inputs = _small_batch(seed=13, count=6)
    bp = BatchProblem(inputs)
    bp.variant = 0 if variant == 16 else 4
    bp.run("mixed16")
    ref = bp.wave.clone()
    bp.variant = variant
    bp.wave.zero_()
    bp.run("mixed16")
    assert torch.equal(bp.wave, ref)
    h0 = bp.waveform(0).cpu().numpy()
    assert oracle.rel_l2(h0, oracle_waveform(inputs[0])) < 1e-6


def test_streamed_full1_vs_fp64(gpu):
    """full1 (3843 modes, 3.16e6 samples): variant 16 equals variant 0 bit for bit
    and stays within the mixed bar of the FP64 GPU path on every sample."""
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    inp = synthetic.make_workload("full1")
    p = ModeSumProblem(inp)
    h64 = p.run("fp64").clone()
    p.variant = 0
    h0 = p.run("mixed16").clone()
    p.variant = 16
    h16 = p.run("mixed16")
    assert torch.equal(h16, h0)
    err = oracle.rel_l2(h16.cpu().numpy(), h64.cpu().numpy())
    print(f"full1 streamed rel-L2 vs fp64 {err:.3e}")
    assert err < 1e-6


def test_every_mixed16_variant_same_bits(gpu):
    """All 16 variants of the split-FP16 sum (waves per workgroup, wave
    priority, contraction form, persistent grid size) give the same samples
    bit for bit, single waveform and batch; an unknown variant bit is refused."""
    from fastemriwaveforms_amd import autotune, synthetic
    from fastemriwaveforms_amd._lib import FewRocmError
    from fastemriwaveforms_amd.pipeline import BatchProblem
    from test_gpu_batch import _small_batch
    assert len(autotune.MIXED16_VARIANTS) == 16
    inp = synthetic.make_workload("cpu0")
    ref, _ = _single(inp, 0)
    for v in autotune.MIXED16_VARIANTS:
        got, _ = _single(inp, v)
        assert np.array_equal(got, ref), v
    bp = BatchProblem(_small_batch(seed=17, count=4))
    bp.variant = 0
    bp.run("mixed16")
    bref = bp.wave.clone()
    for v in autotune.MIXED16_VARIANTS:
        bp.variant = v
        bp.run("mixed16")
        assert torch.equal(bp.wave, bref), v
    with pytest.raises(FewRocmError):
        _single(inp, 64)

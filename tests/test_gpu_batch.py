"""GPU tests of the batched mode sum (uneven waveforms in one launch per stage)."""

import numpy as np
import pytest

import oracle
from helpers import oracle_waveform

pytestmark = pytest.mark.gpu


def _small_batch(seed=5, count=7):
    from fastemriwaveforms_amd import synthetic
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        T = float(rng.uniform(0.001, 0.004))
        L = int(rng.integers(8, 40))
        K = int(rng.integers(5, 300))
        out.append(synthetic.make_input(T, 10.0, L, "m_nonneg_30", K, int(rng.integers(1 << 30)),
                                        name=f"b{b}"))
    # edge members: a single mode, and one whose later knots lie past the last sample
    out.append(synthetic.make_input(0.002, 10.0, 10, "m_nonneg_30", 1, 77, name="one_mode"))
    cut = synthetic.make_input(0.003, 10.0, 14, "m_nonneg_30", 40, 78, name="cut")
    cut.num_pts -= 3000
    out.append(cut)
    return out


def test_batch_fp64_and_mixed16_vs_oracle(gpu):
    from fastemriwaveforms_amd.pipeline import BatchProblem
    inputs = _small_batch()
    bp = BatchProblem(inputs)
    assert bp.supported16
    bp.run("fp64")
    h64 = [bp.waveform(b).cpu().numpy() for b in range(bp.B)]
    bp.run("mixed16")
    h16 = [bp.waveform(b).cpu().numpy() for b in range(bp.B)]
    for b, inp in enumerate(inputs):
        ref = oracle_waveform(inp)
        e64 = oracle.rel_l2(h64[b], ref)
        e16 = oracle.rel_l2(h16[b], ref)
        print(f"{inp.name}: K={inp.K} L={inp.L} N={inp.num_pts} fp64 {e64:.2e} mixed16 {e16:.2e}")
        assert e64 <= 1e-12, (inp.name, e64)
        assert e16 < 1e-6, (inp.name, e16)


def test_batch_splines_bitwise(gpu):
    """The batched spline build (each waveform's knot matrix eliminated once,
    every system sweeping its own right-hand side) equals the oracle's dgtsv
    bit for bit: random knots, knots whose spacing ratios force row
    interchanges, and 3 knots, where the matrix is singular and LAPACK's early
    return makes each system run on its own."""
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.pipeline import BATCH_STAGE_SPLINES, BatchProblem
    inputs = _small_batch(seed=21, count=3)[:3]
    rng = np.random.default_rng(8)
    piv = synthetic.make_input(0.003, 10.0, 40, "m_nonneg_30", 60, 81, name="pivots")
    gaps = np.exp(rng.uniform(-6, 6, piv.L - 1))
    piv.t = np.concatenate([[0.0], np.cumsum(gaps)]) * (piv.t[-1] / np.sum(gaps))
    inputs.append(piv)
    inputs.append(synthetic.make_input(0.002, 10.0, 3, "m_nonneg_30", 25, 82, name="three_knots"))
    bp = BatchProblem(inputs)
    bp.stages("fp64", BATCH_STAGE_SPLINES)
    interp = bp.interp.cpu().numpy()
    pcoef = bp.pcoef.cpu().numpy()
    for b, x in enumerate(inputs):
        y_all = np.concatenate([x.teuk.T.real, x.teuk.T.imag])
        ci, _, info = oracle.interpolate_arrays(x.t, y_all)
        assert info == (3 if x.L == 3 else 0), x.name
        ref_i = np.transpose(ci, (0, 2, 1)).ravel()
        cp, _, _ = oracle.interpolate_arrays(x.t, x.phase_knots)
        ref_p = oracle.cubic_to_dense(cp, x.t).ravel()
        oi, op = int(bp.offsets[b, 4]), int(bp.offsets[b, 5])
        assert np.array_equal(interp[oi:oi + ref_i.size], ref_i), x.name
        assert np.array_equal(pcoef[op:op + ref_p.size], ref_p), x.name


def test_batch_fp64_matches_single(gpu):
    """The batched FP64 control run computes each waveform exactly as the
    single-waveform path does (same kernel body, same plan): bit-identical."""
    from fastemriwaveforms_amd.pipeline import BatchProblem, ModeSumProblem
    inputs = _small_batch(seed=9, count=4)
    bp = BatchProblem(inputs)
    bp.run("fp64")
    for b, inp in enumerate(inputs):
        single = ModeSumProblem(inp).run("fp64").cpu().numpy()
        assert np.array_equal(bp.waveform(b).cpu().numpy(), single), inp.name


def test_batch_graph_replays(gpu):
    """hipGraph capture of the whole batch pipeline: every replay, with eager
    work and synchronisations in between, reproduces the eager result."""
    import torch
    from fastemriwaveforms_amd.pipeline import BatchProblem
    inputs = _small_batch(seed=13, count=5)
    bp = BatchProblem(inputs)
    eager = bp.run("mixed16").clone()
    g = bp.capture("mixed16")
    for rep in range(4):
        bp.wave.fill_(1.0 + rep)          # eager work between replays
        junk = torch.randn(1 << 16, device=bp.wave.device).sum()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(bp.wave, eager), rep
        assert torch.isfinite(junk)


def test_batch_autotune(gpu, tmp_path, monkeypatch):
    from fastemriwaveforms_amd import autotune
    from fastemriwaveforms_amd.pipeline import BatchProblem
    monkeypatch.setenv("FEW_AMD_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    bp = BatchProblem(_small_batch(seed=21, count=3))
    best = autotune.tune_batch(bp, reps=1)
    assert best["variant"] in autotune.BATCH_VARIANTS and best["ms"] > 0
    again = autotune.tune_batch(bp, reps=1)  # served from the cache
    assert again == best
    ref = oracle_waveform(bp.inputs[0])
    bp.run("mixed16")
    assert oracle.rel_l2(bp.waveform(0).cpu().numpy(), ref) < 1e-6


def test_single_autotune(gpu, tmp_path, monkeypatch):
    from fastemriwaveforms_amd import autotune, synthetic
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    monkeypatch.setenv("FEW_AMD_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    inp = synthetic.make_workload("cpu0")
    p = ModeSumProblem(inp)
    best = autotune.tune_problem(p, reps=1)
    h = p.run(best["precision"]).cpu().numpy()
    ref = oracle_waveform(inp)
    assert oracle.rel_l2(h, ref) < 1e-6


def test_batch_sum_must_match_plan(gpu):
    """The batch's interval plan depends on the precision and the variant's tile
    shape: a sum-only call with another variant or precision is refused."""
    import torch
    from fastemriwaveforms_amd._lib import FewRocmError
    from fastemriwaveforms_amd.pipeline import (BATCH_STAGE_PLAN, BATCH_STAGE_SPLINES,
                                                BATCH_STAGE_SUM, BatchProblem)
    bp = BatchProblem(_small_batch(seed=9, count=3))
    bp.variant = 0
    ref = bp.run("mixed16").clone()
    bp.variant = 4
    with pytest.raises(FewRocmError, match="prepared"):
        bp.stages("mixed16", BATCH_STAGE_SUM)
    with pytest.raises(FewRocmError, match="prepared"):
        bp.stages("fp64", BATCH_STAGE_SUM)
    bp.wave.zero_()
    bp.stages("mixed16", BATCH_STAGE_SPLINES | BATCH_STAGE_PLAN)
    bp.stages("mixed16", BATCH_STAGE_SUM)
    assert torch.equal(bp.wave, ref)


@pytest.mark.parametrize("prec", ["fp64", "mixed16"])
def test_batch_overwrite_equals_zero_then_accumulate(gpu, prec):
    """Batch stage flag 8 (overwrite with the sum): every sample of every
    waveform written, whatever the buffer held -- bit for bit zeroing and
    accumulating; the batch holds a member whose output ends before its
    trajectory and one with samples after its last knot."""
    import torch
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.pipeline import (BATCH_STAGE_OVERWRITE, BATCH_STAGE_PLAN,
                                                BATCH_STAGE_SPLINES, BATCH_STAGE_SUM,
                                                BatchProblem)
    inputs = _small_batch(seed=31, count=4)
    longer = synthetic.make_input(0.002, 10.0, 12, "m_nonneg_30", 25, 79, name="after")
    longer.num_pts += 2500
    inputs.append(longer)
    bp = BatchProblem(inputs)
    bp.wave.zero_()
    bp.stages(prec, BATCH_STAGE_SPLINES | BATCH_STAGE_PLAN | BATCH_STAGE_SUM)
    ref = bp.wave.clone()
    assert torch.count_nonzero(bp.waveform(bp.B - 1)[-2000:]) == 0
    bp.wave.fill_(complex(float("nan"), 3.0))
    bp.stages(prec, BATCH_STAGE_SUM | BATCH_STAGE_OVERWRITE)
    assert torch.equal(torch.view_as_real(bp.wave), torch.view_as_real(ref))
    bp.wave.fill_(complex(-7.0, float("nan")))
    bp.run(prec)  # every stage, overwriting
    assert torch.equal(torch.view_as_real(bp.wave), torch.view_as_real(ref))


def test_batch_refuses_sum_after_splines_without_plan(gpu):
    """Stage 1 reuses the mixed16 tile region as spline scratch: a sum in the
    same call without stage 2, or a later sum-only call, is refused."""
    from fastemriwaveforms_amd._lib import FewRocmError
    from fastemriwaveforms_amd.pipeline import (BATCH_STAGE_SPLINES, BATCH_STAGE_SUM,
                                                BatchProblem)
    bp = BatchProblem(_small_batch(seed=3, count=2))
    bp.run("mixed16")
    with pytest.raises(FewRocmError, match="stage 2"):
        bp.stages("mixed16", BATCH_STAGE_SPLINES | BATCH_STAGE_SUM)
    bp.stages("mixed16", BATCH_STAGE_SPLINES)
    with pytest.raises(FewRocmError, match="prepared"):
        bp.stages("mixed16", BATCH_STAGE_SUM)


def test_batch_many_intervals(gpu):
    """A batch with more than 1024 knot intervals (the coefficient tiles are then
    prepared one workgroup per interval): mixed16 against the FP64 control on
    every sample and against the oracle on two waveforms."""
    import torch
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.pipeline import BatchProblem
    rng = np.random.default_rng(41)
    inputs = [synthetic.make_input(0.004, 10.0, 100, "m_nonneg_30", int(rng.integers(100, 400)),
                                   900 + b) for b in range(12)]
    bp = BatchProblem(inputs)
    assert int(bp.totals[8]) >= 1024
    h64 = bp.run("fp64").clone()
    h16 = bp.run("mixed16")
    err = float(torch.linalg.vector_norm(h16 - h64) / torch.linalg.vector_norm(h64))
    print(f"many intervals ({int(bp.totals[8])}): mixed16 vs fp64 {err:.3e}")
    assert err < 1e-6
    for b in (0, 11):
        ref = oracle_waveform(inputs[b])
        assert oracle.rel_l2(bp.waveform(b).cpu().numpy(), ref) < 1e-6


def test_batch_256_uneven_full_size(gpu):
    """The 256-waveform uneven batch at its real size (durations U(0.1, 2) yr at
    dt = 10 s, knots U{50..200}, U{50..3000} of the 3843 modes; 8.3e8 samples):
    mixed16 against the FP64 control on every sample, and 8 waveforms spread over
    the batch against the oracle on spread samples."""
    import torch
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.pipeline import BatchProblem
    from helpers import spread_indices
    inputs = synthetic.make_batch(256)
    bp = BatchProblem(inputs)
    assert bp.supported16 and bp.samples > 8e8
    h16 = bp.run("mixed16").clone()
    h64 = bp.run("fp64")
    err = float(torch.linalg.vector_norm(h16 - h64) / torch.linalg.vector_norm(h64))
    print(f"256 uneven waveforms: {bp.samples} samples, mixed16 vs fp64 (all samples) {err:.3e}")
    assert err < 1e-6
    for b in np.linspace(0, 255, 8).astype(int):
        x = inputs[b]
        idx = spread_indices(x.num_pts, count=200, seed=int(b))
        o = int(bp.offsets[b, 6]) // 2
        got = h16[o:o + x.num_pts][torch.from_numpy(idx).to(h16.device)].cpu().numpy()
        e = oracle.rel_l2(got, oracle_waveform(x, idx))
        print(f"  waveform {b}: K={x.K} L={x.L} N={x.num_pts} rel-L2 vs oracle {e:.3e}")
        assert e < 1e-6
    del h16
    torch.cuda.empty_cache()

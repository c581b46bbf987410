"""GPU accuracy of the mixed-precision factorised (MFMA) mode sum.

Bar (the headline metric's): rel-L2 < 1e-6 against the FP64 oracle; the
factorised sum is also compared with the FP64 GPU path on every sample.
"""

import numpy as np
import pytest
import torch

import oracle
from helpers import oracle_waveform, spread_indices

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-6


def _run(inp, precision, variant=0):
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    p = ModeSumProblem(inp)
    p.variant = variant
    return p.run(precision).cpu().numpy(), p


PRECS = ["mixed", "mixed16"]


@pytest.mark.parametrize("prec", PRECS)
def test_mixed_small_full_oracle(gpu, prec):
    from fastemriwaveforms_amd import synthetic
    inp = synthetic.small_input(seed=31, L=16, K=50, T_years=0.004)
    h, p = _run(inp, prec)
    assert p.plan.supported
    ref = oracle_waveform(inp)
    err = oracle.rel_l2(h, ref)
    print(f"small {prec} rel-L2 {err:.3e}")
    assert err < MIXED_TOL


@pytest.mark.parametrize("prec,variant", [("mixed", 0), ("mixed", 4), ("mixed16", 0), ("mixed16", 4)])
def test_mixed_cpu0_vs_oracle(gpu, prec, variant):
    from fastemriwaveforms_amd import synthetic
    inp = synthetic.make_workload("cpu0")
    h, _ = _run(inp, prec, variant)
    ref = oracle_waveform(inp)
    err = oracle.rel_l2(h, ref)
    mm = oracle.mismatch(h, ref)
    print(f"cpu0 {prec} variant {variant}: rel-L2 {err:.3e} mismatch {mm:.3e}")
    assert err < MIXED_TOL


@pytest.mark.parametrize("prec", PRECS)
def test_mixed_full1_vs_fp64(gpu, prec):
    """Workload full1 (3843 modes, 3.16e6 samples): every sample against the FP64
    GPU path (itself within 1e-12 of the oracle), plus spread oracle samples."""
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    inp = synthetic.make_workload("full1")
    p = ModeSumProblem(inp)
    h64 = p.run("fp64").cpu().numpy()
    hm = p.run(prec).cpu().numpy()
    err = oracle.rel_l2(hm, h64)
    idx = spread_indices(inp.num_pts, count=1500)
    ref = oracle_waveform(inp, idx)
    err_o = oracle.rel_l2(hm[idx], ref)
    print(f"full1 {prec}: rel-L2 vs fp64 GPU {err:.3e}, vs oracle samples {err_o:.3e}")
    assert err < MIXED_TOL and err_o < MIXED_TOL


@pytest.mark.parametrize("prec", PRECS)
def test_mixed_explicit_pm_modes(gpu, prec):
    """Workload mixed2 shape (|m| <= l, n in [-55, 55], Y- = 0) on a short duration.
    Its FP16 tile (21 rows x 111 columns) fits LDS only without row padding."""
    from fastemriwaveforms_amd import synthetic
    inp = synthetic.make_input(0.05, 5.0, 60, "m_all_55", 2000, 1002, minus_m_ylm=False)
    h, p = _run(inp, prec)
    assert p.plan.supported16
    assert p.plan.cols == 111
    assert p.plan.folded and p.plan.tile_cols == 56  # columns n and -n share a tile column
    idx = spread_indices(inp.num_pts, count=1500)
    ref = oracle_waveform(inp, idx)
    err = oracle.rel_l2(h[idx], ref)
    print(f"mixed2-shape {prec} rel-L2 {err:.3e}")
    assert err < MIXED_TOL


@pytest.mark.parametrize("case", ["asymmetric_folded", "narrow_unfolded"])
def test_column_folding(gpu, case):
    """The coefficient tiles fold column -n onto n when the n range reaches below
    zero and folding shortens them.  asymmetric_folded: m = 0 modes with n in
    [-30, 10] beside m != 0 modes with |n| <= 5 (n range [-30, 10]; tile columns
    11..30 hold only their -n partner); narrow_unfolded: n in [0, 7] (no negative
    n, no fold).  mixed16 and mixed against the oracle and the FP64 GPU path,
    fact64 against the FP64 GPU path."""
    from fastemriwaveforms_amd import synthetic
    inp = synthetic.small_input(seed=37, L=16, K=60, T_years=0.004)
    rng = np.random.default_rng(5)
    K = len(inp.m)
    if case == "asymmetric_folded":
        zero = np.arange(K) < K // 2
        inp.m = np.where(zero, 0, inp.m).astype(inp.m.dtype)
        inp.n = np.where(zero, rng.integers(-30, 11, K), rng.integers(-5, 6, K)).astype(inp.n.dtype)
        inp.n[0] = -30
    else:
        inp.m = np.zeros_like(inp.m)
        inp.n = rng.integers(0, 8, K).astype(inp.n.dtype)
    h64, p = _run(inp, "fp64")
    assert p.plan.supported and p.plan.supported16 and p.plan.supported64
    if case == "asymmetric_folded":
        assert p.plan.n_lo == -30 and p.plan.cols == 41 and p.plan.folded and p.plan.tile_cols == 31
    else:
        assert not p.plan.folded and p.plan.tile_cols == p.plan.cols
    ref = oracle_waveform(inp)
    for prec in ("mixed16", "mixed"):
        h, _ = _run(inp, prec)
        err, err64 = oracle.rel_l2(h, ref), oracle.rel_l2(h, h64)
        print(f"{case} {prec}: cols {p.plan.cols} -> {p.plan.tile_cols}, rel-L2 vs oracle "
              f"{err:.3e}, vs fp64 GPU {err64:.3e}")
        assert err < MIXED_TOL and err64 < MIXED_TOL
    hf, _ = _run(inp, "fact64")
    errf = oracle.rel_l2(hf, h64)
    print(f"{case} fact64: rel-L2 vs fp64 GPU {errf:.3e}")
    assert errf < 1e-9


@pytest.mark.parametrize("prec", PRECS)
def test_mixed_accumulates_and_edges(gpu, prec):
    """Accumulate-into semantics and interval edges (knots on the sample grid,
    a knot past the end) for the factorised path."""
    import torch
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    inp = synthetic.small_input(seed=23, L=14, K=30, T_years=0.002)
    inp.t[1:-1] = np.round(inp.t[1:-1] / inp.dt) * inp.dt
    inp.num_pts -= 1000
    p = ModeSumProblem(inp)
    init = (torch.arange(inp.num_pts, dtype=torch.float64, device=p.wave.device) * 1e-3).to(torch.complex128)
    p.wave.copy_(init)
    p.build_splines()
    p.mode_sum(prec)
    h = (p.wave - init).cpu().numpy()
    ref = oracle_waveform(inp)
    assert oracle.rel_l2(h, ref) < MIXED_TOL


@pytest.mark.parametrize("prec", ["fp64", "mixed", "mixed16"])
@pytest.mark.parametrize("tail", ["after_trajectory", "output_cut"])
def test_overwrite_equals_zero_then_accumulate(gpu, prec, tail):
    """stage flag 4 (overwrite): every output sample written -- h inside the
    intervals, 0 in the tail after the last knot -- bit for bit what zeroing
    the buffer and accumulating gives, whatever the buffer held before; also
    when the output ends several knot spacings before the trajectory (the
    find_start_inds stop: the intervals after it own no samples)."""
    import torch
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    inp = synthetic.small_input(seed=29, L=14, K=30, T_years=0.002)
    if tail == "after_trajectory":
        inp.num_pts += 3000  # samples after the last knot
    else:
        # the output ends between knots L-4 and L-3: the last two intervals own no samples
        inp.num_pts = int(0.5 * (inp.t[-4] + inp.t[-3]) / inp.dt)
        assert inp.t[-4] < (inp.num_pts - 1) * inp.dt < inp.t[-3]
    p = ModeSumProblem(inp)
    p.build_splines()
    p.wave.zero_()
    p.mode_sum(prec)
    ref = p.wave.clone()
    if tail == "after_trajectory":
        assert torch.count_nonzero(ref[-2000:]) == 0
    else:
        assert torch.count_nonzero(ref[-100:]) == 100
    gen = torch.Generator(device=p.wave.device).manual_seed(5)
    junk = torch.randn(inp.num_pts, 2, dtype=torch.float64, device=p.wave.device, generator=gen)
    p.wave.copy_(torch.view_as_complex(junk))
    p.mode_sum(prec, stage=2, overwrite=True)
    assert torch.equal(p.wave, ref)


def test_mixed16_2yr_full_size(gpu):
    """The 2 yr waveform at its real size (T = 2 yr at dt = 5 s: 12,623,260 samples; 200
    knots; 2000 modes with n in [-55, 55]): the split-FP16 sum against the FP64
    GPU path on every sample and against the oracle on spread samples."""
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    inp = synthetic.make_workload("mixed2")
    assert inp.num_pts == 12623260 and inp.L == 200 and inp.K == 2000
    p = ModeSumProblem(inp)
    assert p.plan.supported16
    h64 = p.run("fp64").cpu().numpy()
    hm = p.run("mixed16").cpu().numpy()
    err = oracle.rel_l2(hm, h64)
    idx = spread_indices(inp.num_pts, count=400)
    ref = oracle_waveform(inp, idx)
    err_o = oracle.rel_l2(hm[idx], ref)
    err_64 = oracle.rel_l2(h64[idx], ref)
    print(f"2 yr at 5 s, 2000 modes, mixed16: rel-L2 vs fp64 GPU {err:.3e}, vs oracle {err_o:.3e}; fp64 vs oracle {err_64:.3e}")
    assert err < MIXED_TOL and err_o < MIXED_TOL and err_64 < 1e-12


def test_stage2_must_match_stage1(gpu):
    """The interval plan and coefficient tiles a stage-1 call leaves in the
    workspace depend on the precision and the variant's tile shape: a stage-2
    call that does not match is refused instead of summing with a stale plan."""
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd._lib import FewRocmError
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    inp = synthetic.small_input(seed=44, L=16, K=40, T_years=0.004)
    p = ModeSumProblem(inp)
    p.build_splines()
    p.variant = 0
    p.mode_sum("mixed16", stage=1)
    p.mode_sum("mixed16", stage=2, overwrite=True)
    ref = p.wave.clone()
    p.variant = 4
    with pytest.raises(FewRocmError, match="prepared"):
        p.mode_sum("mixed16", stage=2, overwrite=True)
    with pytest.raises(FewRocmError, match="prepared"):
        p.mode_sum("mixed", stage=2, overwrite=True)
    p.mode_sum("mixed16", stage=1)
    p.mode_sum("mixed16", stage=2, overwrite=True)
    assert torch.equal(p.wave, ref)
    # another problem's plan on this workspace
    q = ModeSumProblem(inp)
    q.ws = p.ws
    with pytest.raises(FewRocmError, match="another plan"):
        q.mode_sum("mixed16", stage=2, overwrite=True)


def test_single_step_graph_replays(gpu):
    """The captured step (spline build, plan + tiles, sum writing every sample)
    replayed several times, with eager work and synchronisations in between:
    every replay equals the eager step bit for bit."""
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    inp = synthetic.small_input(seed=47, L=30, K=200, T_years=0.01)
    p = ModeSumProblem(inp)
    ref = p.run("mixed16").clone()
    g = p.capture("mixed16")
    for rep in range(4):
        p.wave.fill_(float(rep + 1))
        p.mode_sum("fp64", stage=3)           # eager work between the replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(p.wave, ref), f"replay {rep}"

"""The allocation-free mode-sum entry point, few_get_waveform_plan_noalloc,
against few_get_waveform_plan_async with a caller-owned workspace, through the
C ABI: identical outputs (same kernels), on two calls in a row where the
second has a smaller input, for the FP64 and the split-FP16 sums, with the
overwriting and the accumulating sum and with separate prepare / sum calls."""

import numpy as np
import pytest
import torch

from fastemriwaveforms_amd import _lib, synthetic
from fastemriwaveforms_amd.pipeline import PRECISION_CODES, ModePlan, ModeSumProblem

pytestmark = pytest.mark.gpu


def _stream():
    return int(torch.cuda.current_stream().cuda_stream)


def _noalloc(prob, precision, stage):
    return _lib.lib().few_get_waveform_plan_noalloc(
        prob.wave.data_ptr(), prob.interp.data_ptr(), prob.t.data_ptr(), prob.pcoef.data_ptr(),
        prob.plan.handle, prob.m.data_ptr(), prob.k.data_ptr(), prob.n.data_ptr(), prob.L,
        prob.N, prob.K, prob.ylm.data_ptr(), prob.dt, prob.t.data_ptr(),
        PRECISION_CODES[precision], prob.variant, stage, _stream())


def _same(a, b):
    return torch.equal(torch.view_as_real(a), torch.view_as_real(b))


@pytest.mark.parametrize("precision", ["fp64", "mixed16"])
def test_noalloc_matches_workspace_entry(gpu, precision):
    # the second call is smaller (fewer knots, modes and samples) than the first
    sizes = [(0.02, 40, 300, 5), (0.005, 16, 60, 6)]
    for T, L, K, seed in sizes:
        inp = synthetic.make_input(T, 10.0, L, "m_nonneg_30", K, seed)
        print(f"noalloc {precision}: T={T} L={L} K={K} N={inp.num_pts}", flush=True)
        prob = ModeSumProblem(inp, gpu)
        prob.build_splines()
        start = torch.full_like(prob.wave, complex(0.25, -0.5))

        # overwriting sum (stage 1 + 2 + 4): every sample written
        prob.wave.copy_(start)
        prob.mode_sum(precision, stage=3, overwrite=True)
        ref = prob.wave.clone()
        prob.wave.fill_(complex(float("nan"), float("nan")))
        _lib.check(_noalloc(prob, precision, 7), "get_waveform_plan_noalloc")
        torch.cuda.synchronize()
        assert _same(prob.wave, ref)

        # accumulating sum, prepare and sum as two calls
        prob.wave.copy_(start)
        prob.mode_sum(precision, stage=3)
        ref_acc = prob.wave.clone()
        prob.wave.copy_(start)
        _lib.check(_noalloc(prob, precision, 1), "get_waveform_plan_noalloc (prepare)")
        _lib.check(_noalloc(prob, precision, 2), "get_waveform_plan_noalloc (sum)")
        torch.cuda.synchronize()
        assert _same(prob.wave, ref_acc)
        assert not _same(ref_acc, ref)  # the accumulated start value is there


def test_noalloc_refuses_beyond_scratch(gpu):
    """A mixed16 call whose coefficient tiles exceed the static array is refused
    before anything runs (800 knots of the 3843-mode set: ~36 MB)."""
    _, m, n = synthetic.mode_set("m_nonneg_30")
    k = np.zeros_like(m)
    plan = ModePlan(m, k, n)
    L, N = 800, 1000
    assert plan.workspace_bytes(L, N, "mixed16") > _lib.FEW_NOALLOC_SCRATCH_BYTES
    buf = torch.zeros(16, dtype=torch.float64, device=gpu)
    p = buf.data_ptr()
    rc = _lib.lib().few_get_waveform_plan_noalloc(
        p, p, p, p, plan.handle, p, p, p, L, N, len(m), p, 10.0, p,
        PRECISION_CODES["mixed16"], 0, 3, _stream())
    assert rc != 0
    assert b"scratch array" in _lib.lib().few_last_error()

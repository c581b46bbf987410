"""Differentiable interpolated mode sum: knot values -> waveform, with a
hand-written adjoint behind :class:`torch.autograd.Function`.

Forward (what ``InterpolatedModeSum.sum`` computes,
summation/interpolatedmodesum.py:283-375, from knot values): the amplitude
and phase splines (one launch), then the mode sum in the chosen precision.

Backward: the splines are recomputed from the saved knot values (nothing of
the forward's evaluation is stored), then

1. ``few_waveform_adjoint_async`` -- segmented per-mode reductions over each
   interval's samples: dL/d(amplitude coefficients) and the moments of the
   per-sample phase sensitivities dL/d(phase cubic coefficients);
2. ``few_modes_adjoint_async`` -- the transposed spline build (transposed
   tridiagonal solve) back to dL/d(knot values), amplitudes and phases.

Step 1 runs in FP64 (``backward="fp64"``, the exact derivative of the FP64
sum, per mode on the vector units), as the factorised FP64 moments
(``backward="fact64"``: the per-interval moments sum_i sigma^s conj(g_i) u_r
v_n with FP64 operands and accumulation on the matrix cores, finished per
mode in FP64 -- FP64-accurate gradients at a fraction of the per-mode cost)
or in mixed precision (``backward="mixed16"``: the same moments on
split-FP16 MFMA with FP32 accumulation -- amplitude gradients to ~4e-7, but
the phase gradients, remainders of sums that cancel by up to ~1e4, only to
~1e-4); step 2 is always FP64.  Inputs are real views: amplitudes as (L, K, 2) float64
(``torch.view_as_real`` of the complex (L, K) knot values), phases as (3, L)
float64 (Phi_phi, Phi_theta, Phi_r).  The waveform comes back as (N, 2)
float64; :meth:`DifferentiableModeSum.__call__` wraps it as complex.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .pipeline import PRECISION_CODES, ModePlan, _stream


_BACKWARD_CODES = {"fp64": _lib.FEW_PRECISION_FP64, "fact64": _lib.FEW_PRECISION_FACT64,
                   "mixed16": _lib.FEW_PRECISION_MIXED16}


class DifferentiableModeSum:
    """Mode set, knot times and buffers of one waveform shape (allocated once).

    t: (L,) knot times [s]; m, k, n: (K,) mode indices; ylms: (2K,) complex,
    Y+ then Y- (Y- already times (-1)^l, as InterpolatedModeSum passes it);
    dt: sample spacing; out_len: number of output samples.
    """

    def __init__(self, t, m, k, n, ylms, dt: float, out_len: int, precision: str = "fp64",
                 device: torch.device | str = "cuda", backward: str = "fp64"):
        dev = torch.device(device)
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        f64 = torch.float64
        self.t = torch.as_tensor(np.ascontiguousarray(t, np.float64)).to(dev)
        self.m = torch.as_tensor(np.ascontiguousarray(m, np.int32)).to(dev)
        self.k = torch.as_tensor(np.ascontiguousarray(k, np.int32)).to(dev)
        self.n = torch.as_tensor(np.ascontiguousarray(n, np.int32)).to(dev)
        self.ylm = torch.as_tensor(np.ascontiguousarray(ylms, np.complex128)).to(dev)
        self.L, self.K, self.N = int(self.t.numel()), int(self.m.numel()), int(out_len)
        self.dt = float(dt)
        if precision not in PRECISION_CODES:
            raise ValueError(f"precision must be one of {sorted(PRECISION_CODES)}")
        self.precision = precision
        if backward not in _BACKWARD_CODES:
            raise ValueError(f"backward must be one of {sorted(_BACKWARD_CODES)}")
        self.backward = backward
        self.variant = 0
        L, K = self.L, self.K
        if L < 4:
            raise ValueError("the adjoint needs at least 4 knots (the 3-knot not-a-knot system "
                             "is singular)")
        self.plan = ModePlan(np.asarray(m), np.asarray(k), np.asarray(n))
        if precision == "mixed16" and not self.plan.supported16:
            raise ValueError("this mode set does not fit the mixed16 tile; use 'mixed'")
        if precision == "mixed" and not self.plan.supported:
            raise ValueError("this mode set does not fit the factorised tile; use 'fp64'")
        if precision == "fact64" and not self.plan.supported64:
            raise ValueError("this mode set does not fit the FP64 factorised tile; use 'fp64'")
        self.interp = torch.empty((4, L, 2 * K), dtype=f64, device=dev)
        self.pcoef = torch.empty((8, L - 1, 3), dtype=f64, device=dev)
        self.scratch = torch.empty(4 * L * (2 * K + 3), dtype=f64, device=dev)
        self.g_interp = torch.empty((4, L, 2 * K), dtype=f64, device=dev)
        self.g_phase = torch.empty((4, L, 3), dtype=f64, device=dev)
        fwd = self.plan.workspace_bytes(L, self.N, precision)
        adj = int(lib().few_adjoint_workspace_bytes(L, self.N, K, _lib.FEW_PRECISION_FP64))
        adj16 = 0
        if backward != "fp64":
            adj16 = int(lib().few_plan_adjoint_workspace_bytes(self.plan.handle, L, self.N,
                                                               _BACKWARD_CODES[backward]))
        self.ws = torch.empty(max(fwd, adj, adj16), dtype=torch.uint8, device=dev)

    # -- device stages -------------------------------------------------------
    def prepare(self, teuk_ri: torch.Tensor, phase_knots: torch.Tensor) -> None:
        rc = lib().few_prepare_modes_async(teuk_ri.data_ptr(), phase_knots.data_ptr(),
                                           self.t.data_ptr(), self.L, self.K,
                                           self.interp.data_ptr(), self.pcoef.data_ptr(),
                                           self.scratch.data_ptr(), _stream())
        check(rc, "prepare_modes")

    def sum_into(self, wave: torch.Tensor, precision: str) -> None:
        rc = lib().few_get_waveform_plan_async(
            wave.data_ptr(), self.interp.data_ptr(), self.t.data_ptr(), self.pcoef.data_ptr(),
            self.plan.handle, self.m.data_ptr(), self.k.data_ptr(), self.n.data_ptr(), self.L,
            self.N, self.K, self.ylm.data_ptr(), self.dt, self.t.data_ptr(),
            PRECISION_CODES[precision], self.variant, 3, self.ws.data_ptr(), self.ws.numel(),
            _stream())
        check(rc, "get_waveform")

    def adjoint(self, grad_wave: torch.Tensor, grad_teuk: torch.Tensor,
                grad_phase: torch.Tensor | None) -> None:
        """grad_wave (N,) complex128 -> grad_teuk (L, K, 2), grad_phase (3, L) (or None).
        The splines in self.interp / self.pcoef must be those of the knot values."""
        prec = _BACKWARD_CODES[self.backward]
        rc = lib().few_waveform_adjoint_plan_async(
            grad_wave.data_ptr(), self.interp.data_ptr(), self.t.data_ptr(),
            self.pcoef.data_ptr(), self.plan.handle, self.m.data_ptr(), self.k.data_ptr(),
            self.n.data_ptr(), self.L, self.N, self.K, self.ylm.data_ptr(), self.dt,
            self.t.data_ptr(), self.g_interp.data_ptr(), self.g_phase.data_ptr(), prec,
            self.ws.data_ptr(), self.ws.numel(), _stream())
        check(rc, "waveform_adjoint")
        rc = lib().few_modes_adjoint_async(
            self.g_interp.data_ptr(), self.g_phase.data_ptr() if grad_phase is not None else None,
            self.t.data_ptr(), self.L, self.K, grad_teuk.data_ptr(),
            grad_phase.data_ptr() if grad_phase is not None else None, self.scratch.data_ptr(),
            _stream())
        check(rc, "modes_adjoint")

    # -- autograd ------------------------------------------------------------
    def __call__(self, teuk, phase_knots: torch.Tensor) -> torch.Tensor:
        """teuk: (L, K) complex128 or (L, K, 2) float64; phase_knots (3, L) float64.
        Returns the complex128 waveform (N,), differentiable in both inputs."""
        teuk_ri = torch.view_as_real(teuk) if teuk.is_complex() else teuk
        h = ModeSumFunction.apply(self, teuk_ri, phase_knots)
        return torch.view_as_complex(h)


def _check_input(x: torch.Tensor, shape, name, dev) -> None:
    if x.dtype != torch.float64 or tuple(x.shape) != tuple(shape):
        raise ValueError(f"{name}: expected float64 {tuple(shape)}, got {x.dtype} {tuple(x.shape)}")
    if x.device != dev:
        raise ValueError(f"{name}: expected a tensor on {dev}, got {x.device}")


class ModeSumFunction(torch.autograd.Function):
    """(engine, teuk (L, K, 2), phase_knots (3, L)) -> waveform (N, 2), all float64."""

    @staticmethod
    def forward(ctx, engine: DifferentiableModeSum, teuk_ri: torch.Tensor,
                phase_knots: torch.Tensor) -> torch.Tensor:
        _check_input(teuk_ri, (engine.L, engine.K, 2), "teuk", engine.device)
        _check_input(phase_knots, (3, engine.L), "phase_knots", engine.device)
        teuk_ri = teuk_ri.contiguous()
        phase_knots = phase_knots.contiguous()
        wave = torch.zeros(engine.N, dtype=torch.complex128, device=engine.device)
        engine.prepare(teuk_ri, phase_knots)
        engine.sum_into(wave, engine.precision)
        ctx.engine = engine
        ctx.save_for_backward(teuk_ri, phase_knots)
        return torch.view_as_real(wave)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        engine: DifferentiableModeSum = ctx.engine
        teuk_ri, phase_knots = ctx.saved_tensors
        need_a, need_p = ctx.needs_input_grad[1], ctx.needs_input_grad[2]
        if not (need_a or need_p):
            return None, None, None
        g = torch.view_as_complex(grad_out.to(torch.float64).contiguous())
        engine.prepare(teuk_ri, phase_knots)  # recompute the splines
        grad_teuk = torch.empty_like(teuk_ri)
        grad_phase = torch.empty_like(phase_knots) if need_p else None
        engine.adjoint(g, grad_teuk, grad_phase)
        return None, (grad_teuk if need_a else None), grad_phase

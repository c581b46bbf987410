"""Knot values -> waveform on the GPU, device-resident end to end.

:class:`ModeSumProblem` holds one waveform's inputs (knot times, complex knot
amplitudes, phase knot values, Y_lm, mode indices) in HBM together with every
buffer the path needs, so a step is exactly:

1. one spline launch: 2K amplitude splines + 3 phase splines (the phase
   pieces written in the dense-output layout get_waveform reads),
2. one planning launch (interval -> sample tiles, find_start_inds rule),
3. one mode-sum launch.

No host synchronisation and no allocation inside :meth:`run`, so a step can
be captured in a HIP graph.
"""

from __future__ import annotations

import sys

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .synthetic import ModeSumInput

# fp64: the reference arithmetic (direct sum); fact64: the factorised sum in
# FP64 on the matrix cores; mixed: FP32 MFMA; mixed16: split-FP16 MFMA
PRECISION_CODES = {"fp64": _lib.FEW_PRECISION_FP64, "fact64": _lib.FEW_PRECISION_FACT64,
                   "mixed": _lib.FEW_PRECISION_MIXED, "mixed16": _lib.FEW_PRECISION_MIXED16}


class ModePlan:
    """Host-built, device-uploaded description of a mode set (rows = signed (m, k),
    columns = n, gather lists) used by the factorised sum.  Create once per mode set."""

    def __init__(self, m, k, n):
        self.m = np.ascontiguousarray(m, np.int32)
        self.k = np.ascontiguousarray(k, np.int32)
        self.n = np.ascontiguousarray(n, np.int32)
        self.K = len(self.m)
        self._lib = lib()  # held for __del__, which may run after module teardown
        self.handle = self._lib.few_mode_plan_create(self.m.ctypes.data, self.k.ctypes.data,
                                                 self.n.ctypes.data, self.K)
        if not self.handle:
            check(1, "mode_plan_create")
        info = np.zeros(7, np.int32)
        check(lib().few_mode_plan_info(self.handle, info.ctypes.data), "mode_plan_info")
        self.rows, self.cols, self.n_lo = int(info[1]), int(info[2]), int(info[3])
        self.row_blocks, self.ksteps = int(info[4]), int(info[5])
        self.supported = bool(info[6] & 1)
        self.supported16 = bool(info[6] & 2)
        self.supported64 = bool(info[6] & 4)
        # complex columns of the coefficient tiles (columns n and -n share one when folded)
        self.folded = bool(info[6] & 8)
        self.tile_cols = (max(-self.n_lo, self.n_lo + self.cols - 1) + 1 if self.folded
                          else self.cols)

    def workspace_bytes(self, L, N, precision) -> int:
        return int(lib().few_plan_workspace_bytes(self.handle, L, N, PRECISION_CODES[precision]))

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter exit the process releases the device memory itself
        h = getattr(self, "handle", None)
        if h and not _finalizing():
            self._lib.few_mode_plan_destroy(h)
        self.handle = None


def _stream() -> int:
    return int(torch.cuda.current_stream().cuda_stream)


class ModeSumProblem:
    def __init__(self, inp: ModeSumInput, device: torch.device | str = "cuda"):
        dev = torch.device(device)
        self.inp = inp
        self.L, self.K, self.N = inp.L, inp.K, inp.num_pts
        self.dt = float(inp.dt)
        f64 = torch.float64
        self.t = torch.from_numpy(np.ascontiguousarray(inp.t, np.float64)).to(dev)
        self.teuk = torch.from_numpy(np.ascontiguousarray(inp.teuk, np.complex128)).to(dev)
        self.phase_knots = torch.from_numpy(np.ascontiguousarray(inp.phase_knots, np.float64)).to(dev)
        self.ylm = torch.from_numpy(np.ascontiguousarray(inp.ylms, np.complex128)).to(dev)
        self.m = torch.from_numpy(np.ascontiguousarray(inp.m, np.int32)).to(dev)
        self.k = torch.from_numpy(np.ascontiguousarray(inp.k, np.int32)).to(dev)
        self.n = torch.from_numpy(np.ascontiguousarray(inp.n, np.int32)).to(dev)
        L, K = self.L, self.K
        self.interp = torch.empty((4, L, 2 * K), dtype=f64, device=dev)
        self.pcoef = torch.empty((8, L - 1, 3), dtype=f64, device=dev)
        # phases handed over as dense output (the trajectory's DOPR853 coefficients) are an
        # input: only the amplitudes are splined per step
        self.dense_input = inp.dense is not None
        if self.dense_input:
            self.pcoef.copy_(torch.from_numpy(np.ascontiguousarray(inp.dense, np.float64)))
        self.scratch = torch.empty(4 * L * (2 * K + 3), dtype=f64, device=dev)
        self.wave = torch.zeros(self.N, dtype=torch.complex128, device=dev)
        self.plan = ModePlan(inp.m, inp.k, inp.n)
        nbytes = max(self.plan.workspace_bytes(L, self.N, p) for p in PRECISION_CODES)
        self.ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        self.variant = 0

    @property
    def mode_samples(self) -> int:
        return self.K * self.N

    def build_splines(self) -> None:
        pk = None if self.dense_input else self.phase_knots.data_ptr()
        pc = None if self.dense_input else self.pcoef.data_ptr()
        rc = lib().few_prepare_modes_async(
            torch.view_as_real(self.teuk).data_ptr(), pk, self.t.data_ptr(), self.L, self.K,
            self.interp.data_ptr(), pc, self.scratch.data_ptr(), _stream())
        check(rc, "prepare_modes")

    def mode_sum(self, precision: str = "fp64", stage: int = 3, overwrite: bool = False) -> None:
        """stage: 1 = plan intervals + coefficient tiles, 2 = the summation kernel, 3 = both.
        overwrite (with the summation): write the waveform instead of adding to it -- every
        sample is written, 0 after the last interval -- so no zeroing pass is needed."""
        if overwrite and stage & 2:
            stage |= 4
        rc = lib().few_get_waveform_plan_async(
            self.wave.data_ptr(), self.interp.data_ptr(), self.t.data_ptr(), self.pcoef.data_ptr(),
            self.plan.handle, self.m.data_ptr(), self.k.data_ptr(), self.n.data_ptr(), self.L,
            self.N, self.K, self.ylm.data_ptr(), self.dt, self.t.data_ptr(),
            PRECISION_CODES[precision], self.variant, stage, self.ws.data_ptr(), self.ws.numel(),
            _stream())
        check(rc, "get_waveform")

    def run(self, precision: str = "fp64") -> torch.Tensor:
        """One full step: build splines, plan, sum (writing every output sample)."""
        self.build_splines()
        self.mode_sum(precision, overwrite=True)
        return self.wave

    def capture(self, precision: str = "mixed16") -> "torch.cuda.CUDAGraph":
        """Capture :meth:`run` in a HIP graph (warm-up run first, on a side stream):
        a replay is one step with no host work per launch."""
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self.run(precision)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run(precision)
        return g


BATCH_STAGE_SPLINES = 1
BATCH_STAGE_PLAN = 2
BATCH_STAGE_SUM = 4
BATCH_STAGE_OVERWRITE = 8  # with BATCH_STAGE_SUM: write every output sample, no zeroing pass


class BatchProblem:
    """Many waveforms (uneven knots, mode subsets, durations) in packed HBM buffers.

    One launch per stage covers the whole batch: splines of every waveform,
    interval plans (+ coefficient tiles), and the persistent mode sum over the
    (waveform, interval, tile) sequence.  :meth:`run` allocates nothing and
    never synchronises, so it can be captured in a HIP graph
    (:meth:`capture`).
    """

    def __init__(self, inputs: list, device: torch.device | str = "cuda"):
        dev = torch.device(device)
        self.inputs = inputs
        B = len(inputs)
        self.B = B
        L = np.array([x.L for x in inputs], np.int32)
        K = np.array([x.K for x in inputs], np.int32)
        N = np.array([x.num_pts for x in inputs], np.int32)
        dt = np.array([x.dt for x in inputs], np.float64)
        m = np.ascontiguousarray(np.concatenate([x.m for x in inputs]), np.int32)
        k = np.ascontiguousarray(np.concatenate([x.k for x in inputs]), np.int32)
        n = np.ascontiguousarray(np.concatenate([x.n for x in inputs]), np.int32)
        self._lib = lib()
        self.handle = self._lib.few_batch_create(B, L.ctypes.data, K.ctypes.data, N.ctypes.data,
                                                 dt.ctypes.data, m.ctypes.data, k.ctypes.data,
                                                 n.ctypes.data)
        if not self.handle:
            check(1, "batch_create")
        totals = np.zeros(10, np.int64)
        offsets = np.zeros((B, 7), np.int64)
        check(lib().few_batch_layout(self.handle, totals.ctypes.data, offsets.ctypes.data),
              "batch_layout")
        self.totals, self.offsets = totals, offsets
        self.supported16 = bool(totals[9] & 2)
        # the union (row, column) grid every waveform's sum contracts
        union = ModePlan(m, k, n)
        self.plan_rows, self.plan_cols = union.rows, union.cols
        self.plan_row_blocks, self.plan_ksteps = union.row_blocks, union.ksteps
        self.plan_tile_cols = union.tile_cols
        del union
        self.mode_samples = int(np.sum(K.astype(np.int64) * N))
        self.samples = int(np.sum(N))
        # phases as dense output (every waveform) are an input in pcoef; else splined per step
        dense = [x.dense is not None for x in inputs]
        if any(dense) and not all(dense):
            raise ValueError("a batch takes dense-output phases for every waveform or for none")
        self.dense_input = all(dense) and B > 0
        # pack host buffers (doubles) once, then one copy each to HBM
        host = [np.zeros(int(totals[i]), np.float64) for i in range(4)]
        pc_host = np.zeros(int(totals[5]), np.float64) if self.dense_input else None
        for b, x in enumerate(inputs):
            o = offsets[b]
            host[0][o[0]:o[0] + x.L] = x.t
            host[1][o[1]:o[1] + 2 * x.L * x.K] = np.ascontiguousarray(x.teuk).view(np.float64).ravel()
            host[2][o[2]:o[2] + 3 * x.L] = np.ascontiguousarray(x.phase_knots).ravel()
            host[3][o[3]:o[3] + 4 * x.K] = np.ascontiguousarray(x.ylms).view(np.float64)
            if self.dense_input:
                pc_host[o[5]:o[5] + 24 * (x.L - 1)] = np.ascontiguousarray(x.dense).ravel()
        self.t, self.teuk, self.pk, self.ylm = (torch.from_numpy(h).to(dev) for h in host)
        self.interp = torch.empty(int(totals[4]), dtype=torch.float64, device=dev)
        if self.dense_input:
            self.pcoef = torch.from_numpy(pc_host).to(dev)
        else:
            self.pcoef = torch.empty(int(totals[5]), dtype=torch.float64, device=dev)
        self.wave = torch.zeros(int(totals[6]) // 2, dtype=torch.complex128, device=dev)
        nbytes = int(lib().few_batch_workspace_bytes(self.handle, PRECISION_CODES["mixed16"]))
        self.ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        self.variant = 0

    def waveform(self, b: int) -> torch.Tensor:
        o = int(self.offsets[b, 6]) // 2
        return self.wave[o:o + self.inputs[b].num_pts]

    def stages(self, precision: str = "mixed16", stage: int = 7) -> None:
        rc = lib().few_batch_run_async(
            self.handle, self.t.data_ptr(), self.teuk.data_ptr(),
            None if self.dense_input else self.pk.data_ptr(),
            self.ylm.data_ptr(), self.interp.data_ptr(), self.pcoef.data_ptr(),
            self.wave.data_ptr(), PRECISION_CODES[precision], self.variant, stage,
            self.ws.data_ptr(), self.ws.numel(), _stream())
        check(rc, "batch_run")

    def run(self, precision: str = "mixed16") -> torch.Tensor:
        """Every stage for every waveform; the sum writes every output sample."""
        self.stages(precision, 15)
        return self.wave

    def capture(self, precision: str = "mixed16") -> "torch.cuda.CUDAGraph":
        """Capture :meth:`run` in a HIP graph (warm-up run first, on a side stream)."""
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self.run(precision)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run(precision)
        return g

    def __del__(self, _finalizing=sys.is_finalizing):
        h = getattr(self, "handle", None)
        if h and not _finalizing():
            self._lib.few_batch_destroy(h)
        self.handle = None

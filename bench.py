#!/usr/bin/env python3
"""Benchmark of the interpolated mode sum on one MI355X.

Headline workload (default): the batched uneven workload -- 256 waveforms with
durations U(0.1, 2) yr at dt = 10 s, knots U{50..200} and U{50..3000} modes of
the 3843-mode set (l = 2..10, m = 0..l, n = -30..30), seeded synthetic inputs
resident in HBM.  One step = every waveform's spline build (2K+3 systems each),
interval plans and coefficient tiles, and the persistent mode sum over the
whole (waveform, interval, tile) sequence, which writes every output sample:
the whole path from knot values to 256 complex128 waveforms.  The timed steps
replay the step captured once as a HIP graph (as a serving loop would); the
per-stage breakdown comes from eager steps with HIP events between the
launches, on the stream the kernels run on.

The headline path is the mixed-precision factorised sum (FP64 phases and
spline build; the n-contraction on FP16 MFMA with hi/lo split operands and
FP32 accumulation); the plain FP64 GPU sum (the reference's arithmetic) is
timed beside it as the denominator of the speed-up.  Further legs: the single
full-year waveform (all 3843 modes), the 2 yr waveform at 5 s (2000
modes with n to 55), the differentiation (forward + backward), the
frequency-domain sum, AAK, interp2D.  The oracle (oracle/, test
infrastructure) runs only in the CPU leg: as the CPU baseline and as the
checker of values the GPU legs kept.  Prints ONE JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6, "fp32_mfma": 157.3, "fp16_mfma": 2500.0}  # MI355X TFLOP/s (dense)
# HBM bytes per launch of fact16_kernel on full1, from two rocprofv3 PMC runs
# (FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00, factors from a 1 GiB copy in the same
# process): profiles/fact16_full1_hbm_traffic.md
FACT16_FULL1_TRAFFIC = 2.00 * 11916.8 * 1024 + 1.00 * 49316.8 * 1024
# HBM bytes per launch of batch_fact16_kernel on the 256-waveform batch, from two
# rocprofv3 PMC runs (FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00, factors from a 1 GiB
# copy in the same process)
BATCH_FACT16_TRAFFIC = 2.00 * 712088.5 * 1024 + 1.00 * 13008300.7 * 1024
BATCH_TRAFFIC_SOURCE = "profiles/batch_fact16_256wave_hbm_traffic.md (rocprofv3 PMC)"
DIRECT_FLOP = 37  # algorithmic FLOP per mode-sample of the direct sum (sincos counted 0)


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def spread_indices(N, count=2000, seed=0):
    """Sample indices spread over [0, N): both ends, a uniform grid, random draws."""
    rng = np.random.default_rng(seed)
    parts = [np.arange(min(20, N)), np.arange(max(0, N - 20), N),
             np.linspace(0, N - 1, count // 2).astype(np.int64), rng.integers(0, N, count // 2)]
    return np.unique(np.concatenate(parts)).astype(np.int64)


def cpu_baseline(inp, seconds_target: float = 12.0):
    """The oracle, single thread (the reference CPU backend is single-threaded), on a
    bounded sample of the same workload: evenly spread output samples, all modes."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle
    from helpers import oracle_coefficients
    interp_tl, pcoef = oracle_coefficients(inp)
    args = (interp_tl, inp.t, pcoef, inp.m, inp.k, inp.n, inp.L, inp.num_pts, inp.K, inp.ylms,
            inp.dt, inp.t)
    probe = np.linspace(0, inp.num_pts - 1, 200).astype(np.int64)
    t0 = time.perf_counter()
    oracle.waveform_samples(probe, *args, nthreads=1)
    per = (time.perf_counter() - t0) / len(probe)
    nsamp = int(min(inp.num_pts, max(1000, seconds_target / max(per, 1e-9))))
    idx = np.linspace(0, inp.num_pts - 1, nsamp).astype(np.int64)
    t0 = time.perf_counter()
    oracle.waveform_samples(idx, *args, nthreads=1)
    el = time.perf_counter() - t0
    # the same restatement with OpenMP over samples, one thread per physical core
    nth, model = physical_cores()
    t0 = time.perf_counter()
    oracle.waveform_samples(idx, *args, nthreads=nth)
    el_omp = time.perf_counter() - t0
    return {"value": nsamp * inp.K / el, "unit": "mode-samples/s", "cores": 1, "kind": "port",
            "omp": {"value": nsamp * inp.K / el_omp, "threads": nth, "cpu_model": model},
            "sample": f"{nsamp} evenly spread output samples x {inp.K} modes of the same workload "
                      f"({nsamp * inp.K:.3g} mode-samples, {el:.1f} s); oracle/few_oracle.c "
                      "restating interpolate.cu's CPU branch (-O2 -ffp-contract=off), 1 thread"}


def cpu_checks(inp, head, batch_chk, fd_chk, aak_chk=None, amp_chk=None, mixed2_chk=None):
    """The CPU leg's checker role: the oracle's values at the samples / bins the
    GPU legs kept, compared with them (rel-L2), and the oracle's frequency-domain
    sum timed on the same check (1 thread)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle
    from helpers import oracle_waveform
    nth = physical_cores()[0]
    res = {}
    if head is not None:
        idx, h_idx = head
        res["headline_ref"] = oracle_waveform(inp, idx, nthreads=nth)
        res["headline"] = oracle.rel_l2(h_idx, res["headline_ref"])
    if batch_chk is not None:
        errs = [oracle.rel_l2(vals, oracle_waveform(x, idx, nthreads=nth))
                for x, idx, vals in batch_chk]
        res["batch_max_of_8"] = float(max(errs))
    if mixed2_chk is not None:
        x, idx, vals = mixed2_chk
        res["mixed2"] = oracle.rel_l2(vals, oracle_waveform(x, idx, nthreads=nth))
    if fd_chk is not None:
        c = fd_chk
        t0 = time.perf_counter()
        ref = oracle.fd_waveform(c["interp"], c["t"], c["pcoef"], c["m"], c["k"], c["n"], c["t"],
                                 c["freq"], c["start"], c["end"], c["ylms"], c["zero"])
        el = time.perf_counter() - t0
        res["fd"] = oracle.rel_l2(c["gpu"], ref)
        res["fd_cpu"] = {"value": c["bins"] / el, "unit": "mode-bins/s", "cores": 1,
                         "kind": "port", "sample": f"{c['modes']} modes x their {c['bins']} "
                         f"frequency bins of the same workload ({el:.1f} s); "
                         "oracle/few_oracle_fd.c restating interpolate.cu:1142-1500, 1 thread"}
    if aak_chk is not None:
        c = aak_chk
        p = c["prm"]
        t0 = time.perf_counter()
        ref = oracle.aak_waveform(c["dense"], p["M"], p["S"], p["mu"], p["qS"], p["phiS"],
                                  p["qK"], p["phiK"], p["dist"], c["nmodes"], True, c["t"],
                                  c["nchk"], c["dt"])
        el = time.perf_counter() - t0
        res["aak"] = oracle.rel_l2(c["gpu"], ref)
        res["aak_cpu"] = {"value": c["nchk"] / el, "unit": "samples/s", "cores": 1,
                          "kind": "port", "sample": f"the first {c['nchk']} samples of the same "
                          f"workload ({el:.2f} s); oracle/few_oracle_aak.c restating gpuAAK.cu's "
                          "CPU branch, 1 thread"}
    if amp_chk is not None:
        c = amp_chk
        t0 = time.perf_counter()
        ref = oracle.interp2d(c["tx"], c["ty"], c["c"], c["x"], c["y"])
        el = time.perf_counter() - t0
        res["interp2d"] = bool(np.array_equal(c["gpu"], ref))
        res["interp2d_cpu"] = {"value": ref.size / el, "unit": "evaluations/s", "cores": 1,
                               "kind": "port", "sample": f"{c['c'].shape[0]} of the coefficient "
                               f"sets at all points ({el * 1e3:.1f} ms); oracle/few_oracle_amp.c, "
                               "1 thread"}
    return res


def time_graph(prob, precision, steps, warmup):
    """The step (spline build, plan + tiles, sum writing every sample) captured once
    as a HIP graph and replayed: `warmup` untimed replays, then `steps` timed ones,
    synchronised on both sides.  Returns ms per step (host wall clock)."""
    import torch
    g = prob.capture(precision)
    for _ in range(warmup):
        g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3


def time_path(prob, precision, steps, warmup):
    """Eager steps with HIP events around each launch (the per-kernel breakdown)."""
    import torch
    for _ in range(warmup):
        prob.run(precision)
    torch.cuda.synchronize()
    ev = {k: [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(steps)] for k in ("spline", "prepare", "sum")}
    t0 = time.perf_counter()
    for s in range(steps):
        ev["spline"][s][0].record()
        prob.build_splines()
        ev["spline"][s][1].record()
        ev["prepare"][s][0].record()
        prob.mode_sum(precision, stage=1)
        ev["prepare"][s][1].record()
        ev["sum"][s][0].record()
        prob.mode_sum(precision, stage=2, overwrite=True)
        ev["sum"][s][1].record()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms = {k: float(np.mean([a.elapsed_time(b) for a, b in v])) for k, v in ev.items()}
    return el / steps * 1e3, ms


def batch_leg(dev, steps: int, warmup: int, fp64: bool, count: int = 256, sum_reps: int = 5):
    """The headline: the uneven batch (durations U(0.1, 2) yr, knots U{50..200}, modes
    U{50..3000} of the 3843 set, dt = 10 s), one launch per stage for the whole batch,
    replayed as one captured HIP graph, variant from the autotuner; per-stage times
    from eager launches with HIP events on the launching stream; the FP64 control
    run beside it."""
    import torch
    from fastemriwaveforms_amd import autotune, synthetic
    from fastemriwaveforms_amd.pipeline import (BATCH_STAGE_OVERWRITE, BATCH_STAGE_PLAN,
                                                BATCH_STAGE_SPLINES, BATCH_STAGE_SUM,
                                                BatchProblem)
    t0 = time.perf_counter()
    inputs = synthetic.make_batch(count)
    bp = BatchProblem(inputs, dev)
    log(f"[bench] batch: {count} waveforms, {bp.samples} samples, {bp.mode_samples:.4g} "
        f"mode-samples, {int(bp.totals[8])} intervals, ws {bp.ws.numel() / 2**30:.2f} GiB "
        f"(setup {time.perf_counter() - t0:.1f} s)")
    best = autotune.tune_batch(bp, reps=1)
    log(f"[bench] batch autotune: {best}")
    g = bp.capture("mixed16")
    for _ in range(warmup):
        g.replay()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    a.record()
    for _ in range(steps):
        g.replay()
    b.record()
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - w0) / steps * 1e3
    ms_graph = a.elapsed_time(b) / steps
    # stage breakdown (eager, same launches, events on the launching stream)
    stage_ms = {"spline": [], "plan_tiles": [], "sum": []}
    for _ in range(sum_reps):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record()
        bp.stages("mixed16", BATCH_STAGE_SPLINES)
        ev[1].record()
        bp.stages("mixed16", BATCH_STAGE_PLAN)
        ev[2].record()
        bp.stages("mixed16", BATCH_STAGE_SUM | BATCH_STAGE_OVERWRITE)
        ev[3].record()
        torch.cuda.synchronize()
        stage_ms["spline"].append(ev[0].elapsed_time(ev[1]))
        stage_ms["plan_tiles"].append(ev[1].elapsed_time(ev[2]))
        stage_ms["sum"].append(ev[2].elapsed_time(ev[3]))
    stage_ms = {k: float(np.mean(v)) for k, v in stage_ms.items()}
    h16 = bp.wave.clone()
    # kept for the CPU leg's oracle check: 8 waveforms spread over the batch, spread samples
    chk = []
    for bi in np.linspace(0, count - 1, 8).astype(int):
        inp = inputs[bi]
        idx = spread_indices(inp.num_pts, count=300, seed=int(bi))
        chk.append((inp, idx, bp.waveform(int(bi)).cpu().numpy()[idx]))
    out = {"waveforms": count, "samples": bp.samples, "mode_samples": bp.mode_samples,
           "intervals": int(bp.totals[8]), "variant": best["variant"], "ms_per_step": ms_graph,
           "wall_ms_per_step": wall_ms, "waveforms_per_s": count / (ms_graph * 1e-3),
           "mode_samples_per_s": bp.mode_samples / (ms_graph * 1e-3), "stage_ms": stage_ms,
           "rel_l2_vs_oracle_max_of_8": None, "plan_rows": int(bp.plan_rows),
           "plan_cols": int(bp.plan_cols), "plan_tile_cols": int(bp.plan_tile_cols),
           "plan_row_blocks": int(bp.plan_row_blocks)}
    log(f"[bench] batch mixed16 graph: {ms_graph:.2f} ms/step (wall {wall_ms:.2f}), "
        f"{out['waveforms_per_s']:.1f} waveforms/s, {out['mode_samples_per_s']:.4g} "
        f"mode-samples/s, stages {stage_ms}")
    if fp64:
        reps = 2
        bp.run("fp64")
        torch.cuda.synchronize()
        a.record()
        for _ in range(reps):
            bp.run("fp64")
        b.record()
        torch.cuda.synchronize()
        ms64 = a.elapsed_time(b) / reps
        diff = torch.linalg.vector_norm(h16 - bp.wave) / torch.linalg.vector_norm(bp.wave)
        out["fp64_control"] = {"ms_per_step": ms64, "waveforms_per_s": count / (ms64 * 1e-3),
                               "mode_samples_per_s": bp.mode_samples / (ms64 * 1e-3),
                               "mixed16_rel_l2_vs_fp64_all_samples": float(diff)}
        out["speedup_vs_fp64_control"] = ms64 / ms_graph
        log(f"[bench] batch fp64 control: {ms64:.1f} ms/step, speed-up {ms64 / ms_graph:.2f}x, "
            f"mixed16 rel-L2 vs fp64 (all samples) {float(diff):.3e}")
    del bp, h16, g
    torch.cuda.empty_cache()
    return out, chk, inputs


def cpu_baseline_batch(inputs, seconds_target: float = 15.0, nwave: int = 8):
    """The oracle, single thread (the reference CPU backend is single-threaded), on a
    bounded sample of the batch: evenly spread output samples of 8 waveforms spread
    over the batch, all of each waveform's modes; mode-samples/s over that sample.
    The same sample with OpenMP over samples on every physical core is reported
    beside it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle
    from helpers import oracle_coefficients
    pick = [inputs[i] for i in np.linspace(0, len(inputs) - 1, nwave).astype(int)]
    prepared = []
    for inp in pick:
        interp_tl, pcoef = oracle_coefficients(inp)
        prepared.append((inp, (interp_tl, inp.t, pcoef, inp.m, inp.k, inp.n, inp.L, inp.num_pts,
                               inp.K, inp.ylms, inp.dt, inp.t)))
    # per-sample cost probe, then a per-waveform sample count for ~seconds_target in total
    t0 = time.perf_counter()
    ms_probe = 0
    for inp, args in prepared:
        probe = np.linspace(0, inp.num_pts - 1, 50).astype(np.int64)
        oracle.waveform_samples(probe, *args, nthreads=1)
        ms_probe += 50 * inp.K
    per_ms = (time.perf_counter() - t0) / ms_probe
    budget_ms = seconds_target / max(per_ms, 1e-12)
    ksum = sum(inp.K for inp, _ in prepared)
    nsamp = max(200, int(budget_ms / ksum))
    idxs = [np.linspace(0, inp.num_pts - 1, min(nsamp, inp.num_pts)).astype(np.int64)
            for inp, _ in prepared]
    msamp = sum(len(ix) * inp.K for ix, (inp, _) in zip(idxs, prepared))
    t0 = time.perf_counter()
    for ix, (inp, args) in zip(idxs, prepared):
        oracle.waveform_samples(ix, *args, nthreads=1)
    el = time.perf_counter() - t0
    cores, model = physical_cores()
    t0 = time.perf_counter()
    for ix, (inp, args) in zip(idxs, prepared):
        oracle.waveform_samples(ix, *args, nthreads=cores)
    el_omp = time.perf_counter() - t0
    return {"value": msamp / el, "unit": "mode-samples/s", "cores": 1, "kind": "port",
            "omp": {"value": msamp / el_omp, "threads": cores, "cpu_model": model},
            "sample": f"{len(idxs[0])} evenly spread output samples of each of {nwave} waveforms "
                      f"spread over the batch, all their modes ({msamp:.3g} mode-samples, "
                      f"{el:.1f} s); oracle/few_oracle.c restating interpolate.cu's CPU branch "
                      "(-O2 -ffp-contract=off), 1 thread"}


def physical_cores():
    """(physical cores available to this process, CPU model) from /proc/cpuinfo and
    the affinity mask: OpenMP legs use one thread per physical core."""
    avail = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else \
        list(range(os.cpu_count() or 1))
    model, core_of, cur = "unknown", {}, {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if "processor" in cur:
                        core_of[int(cur["processor"])] = (cur.get("physical id", "0"),
                                                          cur.get("core id", cur["processor"]))
                    cur = {}
                    continue
                k, v = (x.strip() for x in line.split(":", 1))
                cur[k] = v
                if k == "model name":
                    model = v
        if "processor" in cur:
            core_of[int(cur["processor"])] = (cur.get("physical id", "0"),
                                              cur.get("core id", cur["processor"]))
    except OSError:
        pass
    phys = {core_of.get(c, (0, c)) for c in avail}
    return max(1, len(phys)), model


def mixed2_leg(dev, steps: int = 4):
    """Mixed precision at larger scale: T = 2 yr, dt = 5 s (12,623,260 samples),
    200 knots, 2000 of the 12,987 modes (|m| <= l <= 10, n in [-55, 55]); the mixed16 step
    (same stages as the headline) timed, every sample compared with the plain FP64 GPU path
    (rel-L2 and mismatch), spread samples kept for the oracle in the CPU leg."""
    import torch
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    inp = synthetic.make_workload("mixed2")
    prob = ModeSumProblem(inp, dev)
    prec = "mixed16" if prob.plan.supported16 else "mixed"
    K, N = inp.K, inp.num_pts
    ms_step, ms = time_path(prob, prec, steps, 1)
    h = prob.wave.cpu().numpy()
    ms64_step, ms64 = time_path(prob, "fp64", 1, 0)
    h64 = prob.wave.cpu().numpy()
    fa, fb = np.fft.fft(h), np.fft.fft(h64)
    overlap = np.vdot(fa, fb) / np.sqrt(np.vdot(fa, fa) * np.vdot(fb, fb))
    out = {"workload": f"T=2 yr, dt=5 s, N={N} samples, L={inp.L} knots, K={K} modes "
                       f"(n in [-55, 55]), plan rows={prob.plan.rows} cols={prob.plan.cols}",
           "precision": prec, "ms_per_step": ms_step, "kernel_ms": ms,
           "mode_samples_per_s": K * N / (ms_step * 1e-3),
           "fp64_ms_per_step": ms64_step, "speedup_vs_fp64": ms64_step / ms_step,
           "rel_l2_vs_fp64_all_samples": float(np.linalg.norm(h - h64) / np.linalg.norm(h64)),
           "mismatch_vs_fp64": float(1.0 - overlap.real)}
    log(f"[bench] mixed2 {prec}: step {ms_step:.3f} ms ({out['mode_samples_per_s']:.4g} "
        f"mode-samples/s), fp64 {ms64_step:.1f} ms; rel-L2 vs fp64 "
        f"{out['rel_l2_vs_fp64_all_samples']:.3e}, mismatch {out['mismatch_vs_fp64']:.3e}")
    idx = spread_indices(N, count=600, seed=9)
    chk = (inp, idx, h[idx].copy())
    del prob, h, h64, fa, fb
    torch.cuda.empty_cache()
    return out, chk


def fd_leg(dev, reps: int = 3, check_modes: int = 6):
    """Frequency-domain sum (FDInterpolatedModeSum, get_waveform_generic_fd) at the
    full1 sizes -- T = 1 yr, dt = 10 s (3,155,815 bins), 120 knots, 3843 modes --
    with the "rising" synthetic chirp (monotonic mode frequencies, as an inspiral's).
    Timed end to end (splines, bin bookkeeping, the sum, h+/hx) and the sum kernel
    alone; a few modes' bins are kept for the CPU leg's oracle check."""
    import ctypes
    import torch
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd._lib import check, lib
    from fastemriwaveforms_amd.fdinterp import FDInterpolatedModeSum, segment_frequency_inds
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    from fastemriwaveforms_amd.summation import CubicSplineInterpolant
    inp = synthetic.make_input(1.0, 10.0, 120, "m_nonneg_30", None, 1001, name="full1-fd",
                               chirp="rising")
    K, L = inp.K, inp.L
    prob = ModeSumProblem(inp, dev)
    prob.build_splines()
    phase_coeffs = prob.pcoef.permute(1, 2, 0).contiguous()      # (L-1, 3, 8)
    ylms_in = inp.ylms.copy()
    ylms_in[K:] *= (-1.0) ** inp.l
    freqs = tuple(torch.from_numpy(f).to(dev) for f in inp.freqs)
    summ = FDInterpolatedModeSum(output_type="fd")
    call = lambda: summ(prob.t, prob.teuk, ylms_in, prob.t, phase_coeffs, inp.l, inp.m, inp.k,
                        inp.n, T=inp.T_years, dt=inp.dt, frequencies=freqs)
    call()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    a.record()
    for _ in range(reps):
        call()
    b.record()
    torch.cuda.synchronize()
    ms_e2e = a.elapsed_time(b) / reps
    wall_e2e = (time.perf_counter() - t0) / reps * 1e3
    # the sum kernel alone on the same inputs
    freq = summ.frequency
    Nf = len(freq)
    y_all = torch.cat([prob.teuk.real.T, prob.teuk.imag.T, freqs[0][None], freqs[1][None],
                       freqs[2][None]])
    sp = CubicSplineInterpolant(prob.t, y_all)
    interp = sp.interp_array.reshape(sp.reshape_shape).permute(0, 2, 1).contiguous()
    s0, s1, zero = segment_frequency_inds(prob.t, *freqs, inp.m, inp.k, inp.n, freq)
    pcoef = phase_coeffs.permute(2, 1, 0).contiguous()            # (8, 3, L-1)
    yl = prob.ylm
    nb = ctypes.c_size_t(0)
    check(lib().few_fd_workspace_bytes(K, L - 1, ctypes.byref(nb)), "fd workspace")
    ws = torch.empty(nb.value, dtype=torch.uint8, device=dev)
    stream = int(torch.cuda.current_stream().cuda_stream)

    def kernel(wave, st, en):
        check(lib().few_get_waveform_generic_fd_async(
            wave.data_ptr(), interp.data_ptr(), prob.t.data_ptr(), pcoef.data_ptr(),
            prob.m.data_ptr(), prob.k.data_ptr(), prob.n.data_ptr(), K, inp.dt,
            prob.t.data_ptr(), L, Nf, freq.data_ptr(), st.data_ptr(), en.data_ptr(), L - 1,
            yl.data_ptr(), zero, 1, 0, ws.data_ptr(), ws.numel(), stream), "fd sum")

    wave = torch.zeros(Nf, dtype=torch.complex128, device=dev)
    kernel(wave, s0, s1)
    torch.cuda.synchronize()
    a.record()
    for _ in range(reps):
        kernel(wave, s0, s1)
    b.record()
    torch.cuda.synchronize()
    ms_k = a.elapsed_time(b) / reps
    lo = s0.clamp(min=0).long()
    hi = s1.clamp(max=Nf - 1).long()
    bins = int((hi - lo + 1).clamp(min=0).sum())
    out = {"workload": f"full1 sizes in the frequency domain: T={inp.T_years} yr, dt={inp.dt} s, "
                       f"{Nf} bins, L={L} knots, K={K} modes, rising chirp",
           "mode_bins": bins, "sum_kernel_ms": ms_k, "mode_bins_per_s": bins / (ms_k * 1e-3),
           "end_to_end_ms": ms_e2e, "end_to_end_wall_ms": wall_e2e,
           "waveforms_per_s": 1e3 / ms_e2e, "rel_l2_vs_oracle": None}
    log(f"[bench] fd: {bins:.4g} mode-bins, sum kernel {ms_k:.1f} ms "
        f"({out['mode_bins_per_s']:.4g} mode-bins/s), end to end {ms_e2e:.1f} ms")
    # a few modes for the oracle check (the other modes' ranges emptied)
    pick = np.linspace(0, K - 1, check_modes).astype(int)
    st, en = s0.clone(), s1.clone()
    mask = torch.ones(K, dtype=torch.bool, device=dev)
    mask[torch.from_numpy(pick).to(dev)] = False
    st[mask] = 1
    en[mask] = 0
    wave.zero_()
    kernel(wave, st, en)
    torch.cuda.synchronize()
    chk_bins = int((en.clamp(max=Nf - 1).long() - st.clamp(min=0).long() + 1).clamp(min=0).sum())
    chk = {"interp": interp.cpu().numpy(), "t": inp.t, "pcoef": pcoef.cpu().numpy(),
           "m": inp.m, "k": inp.k, "n": inp.n, "freq": freq.cpu().numpy(),
           "start": st.cpu().numpy(), "end": en.cpu().numpy(), "ylms": inp.ylms, "zero": zero,
           "gpu": wave.cpu().numpy(), "bins": chk_bins, "modes": check_modes}
    del prob, summ, interp, wave, ws
    torch.cuda.empty_cache()
    return out, chk


def adjoint_leg(dev, reps: int = 3, batch: int = 16, fd_coords: int = 64):
    """Differentiation at full size (T = 1 yr, 3843 modes, 120 knots): loss Re<d, h> with
    d = h + complex Gaussian noise (sigma = 0.1 rms(h)); forward and backward timed with
    HIP events (backward = spline recompute + adjoint sum + transposed spline solves);
    gradients checked against FP64 central differences on 64 seeded coordinates; gradient
    throughput over 16 such waveforms (forward + backward each)."""
    import torch
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.autograd import DifferentiableModeSum
    inp = synthetic.make_workload("full1")
    out = {"workload": "full1 forward + backward, loss Re<d,h>, d = h + 0.1 rms(h) noise"}
    teuk = torch.from_numpy(np.ascontiguousarray(inp.teuk)).to(dev)
    pk = torch.from_numpy(np.ascontiguousarray(inp.phase_knots)).to(dev)
    rng = np.random.default_rng(2024)
    # (forward, backward) pairs.  Forward: "fp64" = the reference arithmetic (direct
    # sum), "fact64" = the factorised sum in FP64 on the matrix cores, "mixed16" = the
    # split-FP16 factorised sum.  Backward: "fp64" = the per-mode FP64 adjoint (the
    # reference's arithmetic, differentiated); "fact64" = the factorised FP64 moments
    # on the FP64 matrix cores; "mixed16" = the split-FP16 moments (approximate phase
    # gradients)
    pairs = {"fp64": ("fp64", "fp64"), "fp64+fact64": ("fp64", "fact64"),
             "fact64": ("fact64", "fact64"), "mixed16+fact64": ("mixed16", "fact64"),
             "mixed16": ("mixed16", "mixed16")}
    engines = {name: DifferentiableModeSum(inp.t, inp.m, inp.k, inp.n, inp.ylms, inp.dt,
                                           inp.num_pts, precision=f, device=dev, backward=b)
               for name, (f, b) in pairs.items()}
    with torch.no_grad():
        h0 = engines["fp64"](teuk, pk)
        sig = 0.1 * float(torch.sqrt(torch.mean(torch.abs(h0) ** 2)))
        noise = (rng.normal(size=inp.num_pts) + 1j * rng.normal(size=inp.num_pts)) * sig / np.sqrt(2)
        d = h0 + torch.from_numpy(noise).to(dev)
    grads = {}
    for prec, eng in engines.items():
        t_ri = torch.view_as_real(teuk).clone().requires_grad_(True)
        p_k = pk.clone().requires_grad_(True)
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
               torch.cuda.Event(enable_timing=True)) for _ in range(reps + 1)]
        for r in range(reps + 1):
            t_ri.grad = None
            p_k.grad = None
            ev[r][0].record()
            h = torch.view_as_complex(ModeSumFunction_apply(eng, t_ri, p_k))
            loss = torch.sum(d.real * h.real + d.imag * h.imag)
            ev[r][1].record()
            loss.backward()
            ev[r][2].record()
        torch.cuda.synchronize()
        fwd = float(np.mean([a.elapsed_time(b) for a, b, _ in ev[1:]]))
        bwd = float(np.mean([b.elapsed_time(c) for _, b, c in ev[1:]]))
        grads[prec] = (t_ri.grad.detach().clone(), p_k.grad.detach().clone())
        out[prec] = {"fwd_ms": fwd, "bwd_ms": bwd, "bwd_over_fwd": bwd / fwd,
                     "forward_precision": pairs[prec][0], "backward_precision": pairs[prec][1]}
        log(f"[bench] adjoint {prec}: fwd {fwd:.2f} ms, bwd {bwd:.2f} ms, "
            f"bwd/fwd {bwd / fwd:.2f}")
    # FD check: FP64 central differences of the FP64 GPU forward (equal to the oracle to ~4e-16)
    eng = engines["fp64"]
    wave = torch.zeros(inp.num_pts, dtype=torch.complex128, device=dev)

    def loss_at(t_ri, p_k):
        wave.zero_()
        eng.prepare(t_ri, p_k)
        eng.sum_into(wave, "fp64")
        return float(torch.sum(d.real * wave.real + d.imag * wave.imag))

    g_t, g_p = grads["fp64"]
    for name in pairs:
        if name == "fp64":
            continue
        ga = float(torch.linalg.vector_norm(grads[name][0] - g_t) / torch.linalg.vector_norm(g_t))
        gph = float(torch.linalg.vector_norm(grads[name][1] - g_p) / torch.linalg.vector_norm(g_p))
        out[name]["grad_rel_l2_vs_fp64"] = {"amplitude_knots": ga, "phase_knots": gph}
        log(f"[bench] adjoint {name} gradient rel-L2 vs fp64: amplitudes {ga:.2e}, "
            f"phases {gph:.2e}")
    if fd_coords <= 0:
        return out
    base_t = torch.view_as_real(teuk).contiguous()
    amp_scale = float(torch.abs(teuk).max())
    ad, fd, ad_f, ad_64 = [], [], [], []
    g_tf, g_pf = grads["mixed16+fact64"]
    g_t64, g_p64 = grads["fact64"]
    for c in range(fd_coords):
        if c < fd_coords * 3 // 4:   # amplitude knot (linear: step relative to max |A|)
            l, mo, part = int(rng.integers(inp.L)), int(rng.integers(inp.K)), int(rng.integers(2))
            # the loss is linear in the amplitudes, so the central difference has no
            # truncation error at any step: a large one keeps the loss's rounding small
            step = 1e-3 * amp_scale
            tp, tm = base_t.clone(), base_t.clone()
            tp[l, mo, part] += step
            tm[l, mo, part] -= step
            fd.append((loss_at(tp, pk) - loss_at(tm, pk)) / (2 * step))
            ad.append(float(g_t[l, mo, part]))
            ad_f.append(float(g_tf[l, mo, part]))
            ad_64.append(float(g_t64[l, mo, part]))
        else:                        # phase knot (Phi_phi or Phi_r): Richardson of 5-point stencils
            q, l = int(rng.choice([0, 2])), int(rng.integers(inp.L))

            def at(dd):
                pq = pk.clone()
                pq[q, l] += dd
                return loss_at(base_t, pq)

            def d5(st):
                return (-at(2 * st) + 8 * at(st) - 8 * at(-st) + at(-2 * st)) / (12 * st)

            fd.append((16 * d5(5e-4) - d5(1e-3)) / 15)
            ad.append(float(g_p[q, l]))
            ad_f.append(float(g_pf[q, l]))
            ad_64.append(float(g_p64[q, l]))
    ad, fd, ad_f, ad_64 = np.array(ad), np.array(fd), np.array(ad_f), np.array(ad_64)
    per = np.abs(ad - fd) / np.maximum(np.abs(fd), 1e-300)
    n_amp = fd_coords * 3 // 4
    # per coordinate, relative to the scale of its gradient array (a single knot can
    # have a derivative far below that scale, where cancellation leaves |FD| tiny)
    scale = np.concatenate([np.full(n_amp, float(g_t.abs().max())),
                            np.full(fd_coords - n_amp, float(g_p.abs().max()))])
    per_scale = np.abs(ad - fd) / scale
    out["fd_check"] = {"coords": fd_coords, "amplitude_coords": fd_coords * 3 // 4,
                       "phase_coords": fd_coords - fd_coords * 3 // 4,
                       "rel_l2": float(np.linalg.norm(ad - fd) / np.linalg.norm(fd)),
                       "max_rel": float(per.max()), "median_rel": float(np.median(per)),
                       "max_rel_amplitude": float(per[:n_amp].max()),
                       "max_rel_phase": float(per[n_amp:].max()),
                       "max_rel_to_gradient_scale": float(per_scale.max()),
                       "steps": "amplitudes: central difference, step 1e-3 max|A| (the loss is "
                                "linear in them); phases: Richardson extrapolation of 5-point "
                                "stencils at 1e-3 and 5e-4 rad (FP64 rounding of 1e5-1e7 rad "
                                "phases floors a 1e-6 rad central difference at ~1e-3)",
                       "function": "FP64 GPU forward (rel-L2 ~4e-16 vs the oracle)",
                       "mixed16_fact64_rel_l2": float(np.linalg.norm(ad_f - fd) / np.linalg.norm(fd)),
                       "fact64_rel_l2": float(np.linalg.norm(ad_64 - fd) / np.linalg.norm(fd))}
    log(f"[bench] adjoint FD check ({fd_coords} coords): rel-L2 {out['fd_check']['rel_l2']:.2e}, "
        f"max per-coordinate {per.max():.2e}, median {np.median(per):.2e}; mixed16+fact64 "
        f"rel-L2 {out['fd_check']['mixed16_fact64_rel_l2']:.2e}; fact64 pair "
        f"rel-L2 {out['fd_check']['fact64_rel_l2']:.2e}")
    # gradient throughput: 16 full-size waveforms (own knots and amplitudes), fwd + bwd each
    if batch <= 0:
        return out
    batch_inputs = [synthetic.make_workload("full1", seed=5000 + b) for b in range(batch)]
    for prec in pairs:
        eng = engines[prec]
        dev_in = [(torch.from_numpy(x.t).to(dev), torch.view_as_real(torch.from_numpy(
            np.ascontiguousarray(x.teuk)).to(dev)).clone().requires_grad_(True),
            torch.from_numpy(np.ascontiguousarray(x.phase_knots)).to(dev).requires_grad_(True))
            for x in batch_inputs]
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for t_b, t_ri, p_k in dev_in:
            eng.t.copy_(t_b)
            h = torch.view_as_complex(ModeSumFunction_apply(eng, t_ri, p_k))
            torch.sum(d.real * h.real + d.imag * h.imag).backward()
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b)
        out[prec]["batch16_ms"] = ms
        out[prec]["gradients_per_s"] = batch / (ms * 1e-3)
        log(f"[bench] adjoint {prec}: {batch} waveforms fwd+bwd in {ms:.1f} ms "
            f"({batch / (ms * 1e-3):.2f} gradients/s)")
        eng.t.copy_(torch.from_numpy(inp.t).to(dev))
        del dev_in
    return out


def aak_leg(dev, reps: int = 3, nmodes: int = 30):
    """AAK summation (AAKSummation, gpuAAK.cu's get_waveform) at 1 yr, dt = 10 s
    (3,155,815 samples), 100 knots, 30 harmonics, LISA long-wavelength
    response; a slice of the output is kept for the CPU leg's oracle check."""
    import torch
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd._lib import check, lib
    rng = np.random.default_rng(77)
    T, L, dt = 1.0, 100, 10.0
    Ts = T * synthetic.YRSID_SI
    t = np.concatenate([[0.0], np.sort(rng.uniform(0.0, 1.0, L - 2)) * Ts, [Ts]])
    tau = t / Ts
    pk = synthetic.phase_phi(t, Ts, "rising")
    knots = np.stack([12.0 - 4.0 * tau, 0.4 - 0.3 * tau ** 2, 0.7 + 0.05 * tau, pk, 0.97 * pk,
                      0.83 * pk])
    # the trajectory's dense-output coefficients, from GPU cubic splines of the knots
    t_d = torch.from_numpy(t).to(dev)
    knots_d = torch.from_numpy(np.ascontiguousarray(knots)).to(dev)
    dense = torch.empty((8, L - 1, 6), dtype=torch.float64, device=dev)
    scratch = torch.empty(4 * L * 6, dtype=torch.float64, device=dev)
    check(lib().few_spline_build_async(knots_d.data_ptr(), L, 1, t_d.data_ptr(), 0, 1,
                                       dense.data_ptr(), 0, 0, 0, 6, L, 1, scratch.data_ptr(),
                                       int(torch.cuda.current_stream().cuda_stream)), "aak knots")
    n_out = synthetic.num_points(t[0], t[-1], T, dt)
    prm = dict(M=1e6, S=0.5, mu=10.0, qS=0.9, phiS=1.3, qK=0.6, phiK=2.1, dist=1.0)
    wave = torch.zeros(n_out, dtype=torch.complex128, device=dev)
    ws = torch.empty(int(lib().few_aak_workspace_bytes(L)), dtype=torch.uint8, device=dev)
    stream = int(torch.cuda.current_stream().cuda_stream)

    def run():
        check(lib().few_aak_waveform_async(
            wave.data_ptr(), dense.data_ptr(), prm["M"], prm["S"], prm["mu"], prm["qS"],
            prm["phiS"], prm["qK"], prm["phiK"], prm["dist"], nmodes, 1, L, n_out, dt,
            t_d.data_ptr(), ws.data_ptr(), ws.numel(), stream), "aak sum")

    run()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        run()
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    out = {"workload": f"AAK, T={T} yr, dt={dt} s, {n_out} samples, L={L} knots, "
                       f"{nmodes} harmonics, LISA long-wavelength response",
           "ms": ms, "samples_per_s": n_out / (ms * 1e-3),
           "harmonic_samples_per_s": n_out * nmodes / (ms * 1e-3), "rel_l2_vs_oracle": None}
    log(f"[bench] aak: {ms:.2f} ms, {out['samples_per_s']:.4g} samples/s "
        f"({out['harmonic_samples_per_s']:.4g} harmonic-samples/s)")
    nchk = 4000
    chk = {"t": t, "dense": dense.cpu().numpy(), "prm": prm, "nmodes": nmodes, "dt": dt,
           "n_out": n_out, "nchk": nchk, "gpu": wave[:nchk].cpu().numpy()}
    del wave, dense, ws
    return out, chk


def interp2d_leg(dev, reps: int = 20):
    """interp2D at the size of a full mode set: 2 x 3843 coefficient sets on a
    38 x 48 knot grid (1,496 coefficients each), 120 trajectory points."""
    import torch
    from fastemriwaveforms_amd._lib import check, lib
    rng = np.random.default_rng(5)
    tx = np.concatenate([[0.0] * 4, np.sort(rng.uniform(0, 1, 30)), [1.0] * 4])
    ty = np.concatenate([[0.0] * 4, np.sort(rng.uniform(0, 1, 40)), [1.0] * 4])
    nc = (len(tx) - 4) * (len(ty) - 4)
    sets, mx = 2 * 3843, 120
    c = rng.normal(size=(sets, nc))
    x, y = rng.uniform(0, 1, mx), rng.uniform(0, 1, mx)
    d = lambda v: torch.from_numpy(np.ascontiguousarray(v)).to(dev)
    txd, tyd, cd, xd, yd = d(tx), d(ty), d(c), d(x), d(y)
    z = torch.empty(sets * mx, dtype=torch.float64, device=dev)
    ws = torch.empty(int(lib().few_interp2d_workspace_bytes(mx)), dtype=torch.uint8, device=dev)
    stream = int(torch.cuda.current_stream().cuda_stream)

    def run():
        check(lib().few_interp2d_async(z.data_ptr(), txd.data_ptr(), len(tx), tyd.data_ptr(),
                                       len(ty), cd.data_ptr(), 3, 3, xd.data_ptr(), mx,
                                       yd.data_ptr(), mx, sets, nc, ws.data_ptr(), ws.numel(),
                                       stream), "interp2D")

    run()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        run()
    b.record()
    torch.cuda.synchronize()
    us = a.elapsed_time(b) / reps * 1e3
    out = {"workload": f"{sets} coefficient sets x {nc} coefficients, {mx} points, bicubic",
           "us": us, "evaluations_per_s": sets * mx / (us * 1e-6), "bitwise_vs_oracle": None}
    log(f"[bench] interp2D: {us:.1f} us, {out['evaluations_per_s']:.4g} evaluations/s")
    chk = {"tx": tx, "ty": ty, "c": c[:64], "x": x, "y": y, "gpu": z.reshape(sets, mx)[:64].cpu().numpy()}
    return out, chk


def ModeSumFunction_apply(eng, t_ri, p_k):
    from fastemriwaveforms_amd.autograd import ModeSumFunction
    return ModeSumFunction.apply(eng, t_ri, p_k)


def fact16_roofline(kernel, rows, cols, tile_cols, row_blocks, samples, sum_ms, traffic=None,
                    traffic_source=None):
    """Roofline entry of a split-FP16 factorised sum launch: algorithmic FLOP = 8 FLOP x
    4 powers x rows x cols per output sample (the complex n-contraction of the (m, n)
    cell grid) x samples; against the published dense FP16 MFMA peak.  The tiles fold
    column -n onto n (v_{-n} = conj(v_n)) when the n range is symmetric, so the kernel
    executes the contraction over tile_cols columns (8 FLOP each: four real products and
    sums); that count and the issue rate of the split products (3 per FP32-accurate
    product, padded to the MFMA tiles) are reported beside it."""
    flop_per_sample = 8 * 4 * rows * cols
    achieved = flop_per_sample * samples / (sum_ms * 1e-3) / 1e12
    executed_per_sample = 8 * 4 * rows * tile_cols
    executed = executed_per_sample * samples / (sum_ms * 1e-3) / 1e12
    kb = (2 * tile_cols + 31) // 32
    issued_per_sample = 3 * 2 * (16 * row_blocks) * (32 * kb)
    issued = issued_per_sample * samples / (sum_ms * 1e-3) / 1e12
    return {"bound": "mfma", "achieved": achieved, "peak": PEAK["fp16_mfma"], "unit": "TFLOP/s",
            "frac": achieved / PEAK["fp16_mfma"], "traffic": traffic, "kernel": kernel,
            "algorithmic": f"8 FLOP x 4 powers x {rows} rows x {cols} cols = {flop_per_sample} "
                           f"FLOP/sample x {samples} samples per launch",
            "kernel_ms": sum_ms, "traffic_source": traffic_source,
            "executed_contraction": {"tflops": executed, "frac_of_peak": executed / PEAK["fp16_mfma"],
                                     "flop_per_sample": executed_per_sample,
                                     "basis": f"{tile_cols} tile columns ({cols} n columns, "
                                              f"{'-n folded onto n' if tile_cols < cols else 'unfolded'})"},
            "split_product_issue": {"tflops": issued, "frac_of_peak": issued / PEAK["fp16_mfma"],
                                    "flop_per_sample": issued_per_sample,
                                    "basis": "3 FP16 products (lo*hi, hi*lo, hi*hi) per "
                                             "FP32-accurate product, rows and columns padded "
                                             f"to {16 * row_blocks} x {32 * kb} real"}}


def single_leg(dev, args):
    """One full-year waveform: T = 1 yr, dt = 10 s (3,155,815 samples), 120 knots, all
    3843 modes; the step captured as a HIP graph and replayed; the plain FP64 GPU
    path beside it."""
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.pipeline import ModeSumProblem
    inp = synthetic.make_workload("full1")
    prob = ModeSumProblem(inp, dev)
    K, N = inp.K, inp.num_pts
    prec = "mixed16" if prob.plan.supported16 else "mixed"
    ms_eager, ms = time_path(prob, prec, 50, args.single_warmup)
    ms_step = time_graph(prob, prec, args.single_steps, args.single_warmup)
    h = prob.wave.cpu().numpy()
    value = K * N / (ms_step * 1e-3)
    log(f"[bench] full1 {prec}: step {ms_step:.3f} ms (graph replay; eager {ms_eager:.3f}) | "
        f"spline {ms['spline']:.3f} prepare {ms['prepare']:.3f} sum {ms['sum']:.3f} ms | "
        f"{value:.4g} mode-samples/s")
    idx = spread_indices(N, count=1500, seed=5)
    out = {"workload": f"T={inp.T_years} yr, dt={inp.dt} s, N={N} samples, "
                       f"L={inp.L} knots, K={K} modes",
           "precision": prec, "ms_per_step": ms_step, "eager_ms_per_step": ms_eager,
           "kernel_ms": ms, "mode_samples_per_s": value, "waveforms_per_s": 1e3 / ms_step,
           "rel_l2_vs_oracle": None,
           # spline build: 32 B per (array, knot) -- the knot value read, c1..c3 written
           "spline_roofline": {"bound": "hbm", "unit": "GB/s", "peak": 8000.0,
                               "achieved": 32.0 * (2 * K + 3) * inp.L / (ms["spline"] * 1e-3) / 1e9,
                               "algorithmic": f"32 B x {2 * K + 3} arrays x {inp.L} knots",
                               "bound_by": "the elimination and back-substitution chains "
                                           "(latency), not bandwidth"},
           "roofline": fact16_roofline("fact16_kernel", prob.plan.rows, prob.plan.cols,
                                       prob.plan.tile_cols, prob.plan.row_blocks, N, ms["sum"],
                                       FACT16_FULL1_TRAFFIC,
                                       "profiles/fact16_full1_hbm_traffic.md (rocprofv3 PMC)")}
    chk = (idx, h[idx].copy())
    if not args.no_fp64:
        ms64_step, ms64 = time_path(prob, "fp64", 3, 1)
        h64 = prob.wave.cpu().numpy()
        fa, fb = np.fft.fft(h), np.fft.fft(h64)
        overlap = np.vdot(fa, fb) / np.sqrt(np.vdot(fa, fa) * np.vdot(fb, fb))
        out["fp64_plain_gpu"] = {
            "ms_per_step": ms64_step, "mode_samples_per_s": K * N / (ms64_step * 1e-3),
            "kernel_ms": ms64, "rel_l2_vs_oracle": None,
            "mixed16_rel_l2_vs_fp64_all_samples": float(np.linalg.norm(h - h64) / np.linalg.norm(h64)),
            "mixed16_mismatch_vs_fp64": float(1.0 - overlap.real),
            "direct_fp64_roofline": {"pipe": "FP64 VALU", "peak": PEAK["fp64"], "unit": "TFLOP/s",
                                     "achieved": DIRECT_FLOP * K * N / (ms64["sum"] * 1e-3) / 1e12,
                                     "algorithmic": f"{DIRECT_FLOP} FLOP/mode-sample, sincos "
                                                    "counted 0"}}
        out["speedup_vs_fp64_plain_gpu"] = ms64_step / ms_step
        if prob.plan.supported64:
            # the factorised sum in FP64 on the matrix cores (near-FP64 results)
            tuned_variant, prob.variant = prob.variant, 0  # fact64: 8 waves x 16 samples
            ms_f64_step, ms_f64 = time_path(prob, "fact64", 10, 2)
            hf = prob.wave.cpu().numpy()
            prob.variant = tuned_variant
            out["fact64"] = {
                "ms_per_step": ms_f64_step, "kernel_ms": ms_f64,
                "mode_samples_per_s": K * N / (ms_f64_step * 1e-3),
                "rel_l2_vs_fp64_all_samples": float(np.linalg.norm(hf - h64) / np.linalg.norm(h64)),
                "speedup_vs_fp64_plain_gpu": ms64_step / ms_f64_step,
                "roofline": {"pipe": "FP64 MFMA", "peak": PEAK["fp64"], "unit": "TFLOP/s",
                             "achieved": 32.0 * prob.plan.rows * prob.plan.cols * N /
                             (ms_f64["sum"] * 1e-3) / 1e12,
                             "algorithmic": "8 FLOP x 4 powers x rows x cols per sample",
                             # the folded tiles contract over tile_cols columns, so the
                             # algorithmic rate can exceed the peak; the executed one cannot
                             "executed_contraction": {
                                 "tflops": 32.0 * prob.plan.rows * prob.plan.tile_cols * N /
                                 (ms_f64["sum"] * 1e-3) / 1e12,
                                 "basis": f"8 FLOP x 4 powers x rows x {prob.plan.tile_cols} "
                                          "tile columns per sample"}}}
            rf = out["fact64"]["roofline"]
            rf["frac"] = rf["achieved"] / PEAK["fp64"]
            rf["executed_contraction"]["frac_of_peak"] = rf["executed_contraction"]["tflops"] / PEAK["fp64"]
            log(f"[bench] full1 fact64: step {ms_f64_step:.3f} ms (sum {ms_f64['sum']:.3f}), "
                f"{ms64_step / ms_f64_step:.1f}x the FP64 direct step, rel-L2 vs fp64 "
                f"{out['fact64']['rel_l2_vs_fp64_all_samples']:.2e}")
        chk = (idx, h[idx].copy(), h64[idx].copy())
        log(f"[bench] full1 fp64: step {ms64_step:.3f} ms, speed-up {ms64_step / ms_step:.2f}x, "
            f"mixed16 rel-L2 vs fp64 {out['fp64_plain_gpu']['mixed16_rel_l2_vs_fp64_all_samples']:.3e}")
    del prob
    return out, chk, inp


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the headline step (the 256-waveform batch) is ~0.1 s: 5 untimed replays bring the
    # clocks to their loaded state, 20 timed replays ~2 s
    ap.add_argument("--steps", type=int, default=20, help="timed headline (batch) steps")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--single-steps", type=int, default=100,
                    help="timed steps of the single full-year waveform leg")
    ap.add_argument("--single-warmup", type=int, default=50)
    ap.add_argument("--no-single", action="store_true", help="skip the single full-year waveform leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 controls")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the differentiation leg (forward + backward, FD check)")
    ap.add_argument("--no-fd", action="store_true",
                    help="skip the frequency-domain sum leg")
    ap.add_argument("--no-mixed2", action="store_true",
                    help="skip the 2 yr waveform leg (5 s, 200 knots, 2000 modes, n to 55)")
    ap.add_argument("--no-aak", action="store_true",
                    help="skip the AAK and interp2D legs")
    ap.add_argument("--fd-coords", type=int, default=64,
                    help="coordinates of the finite-difference check (0 skips it)")
    ap.add_argument("--grad-batch", type=int, default=16,
                    help="waveforms of the gradient-throughput measurement (0 skips it)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a known-traffic "
                         "kernel that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (the sum would shard "
              "by output time ranges with no exchange). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)

    if args.copy_mb > 0:
        nel = args.copy_mb * (1 << 20) // 8
        src = torch.ones(nel, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy: {args.copy_mb} MiB read + {args.copy_mb} MiB written")
        del src, dst

    batch, batch_chk, batch_inputs = batch_leg(dev, args.steps, args.warmup, not args.no_fp64,
                                               args.batch_count)
    roofline = fact16_roofline("batch_fact16_kernel", batch["plan_rows"], batch["plan_cols"],
                               batch["plan_tile_cols"], batch["plan_row_blocks"], batch["samples"],
                               batch["stage_ms"]["sum"], BATCH_FACT16_TRAFFIC,
                               BATCH_TRAFFIC_SOURCE)
    log(f"[bench] batch sum roofline: {roofline['achieved']:.1f} TFLOP/s = "
        f"{roofline['frac']:.3f} of the dense FP16 peak (executed folded contraction "
        f"{roofline['executed_contraction']['frac_of_peak']:.3f}, split-product issue "
        f"{roofline['split_product_issue']['frac_of_peak']:.3f})")

    single = single_chk = inp1 = None
    if not args.no_single:
        single, single_chk, inp1 = single_leg(dev, args)

    mixed2 = mixed2_chk = None
    if not args.no_mixed2:
        mixed2, mixed2_chk = mixed2_leg(dev)

    fd = fd_chk = None
    if not args.no_fd:
        fd, fd_chk = fd_leg(dev)

    aak = aak_chk = amp = amp_chk = None
    if not args.no_aak:
        aak, aak_chk = aak_leg(dev)
        amp, amp_chk = interp2d_leg(dev)

    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_leg(dev, batch=args.grad_batch, fd_coords=args.fd_coords)

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU leg: oracle checks and baseline (1 thread)")
        head = (single_chk[0], single_chk[1]) if single_chk is not None else None
        chk = cpu_checks(inp1, head, batch_chk, fd_chk, aak_chk, amp_chk, mixed2_chk)
        batch["rel_l2_vs_oracle_max_of_8"] = chk["batch_max_of_8"]
        log(f"[bench] batch rel-L2 (8 waveforms vs oracle) max {chk['batch_max_of_8']:.3e}")
        if single is not None:
            single["rel_l2_vs_oracle"] = chk["headline"]
            log(f"[bench] full1 rel-L2 vs oracle ({len(single_chk[0])} samples) "
                f"{chk['headline']:.3e}")
            if len(single_chk) > 2 and "fp64_plain_gpu" in single:
                ref = chk["headline_ref"]
                single["fp64_plain_gpu"]["rel_l2_vs_oracle"] = float(
                    np.linalg.norm(single_chk[2] - ref) / np.linalg.norm(ref))
        if mixed2 is not None:
            mixed2["rel_l2_vs_oracle_600_samples"] = chk["mixed2"]
            log(f"[bench] mixed2 rel-L2 vs oracle (600 samples) {chk['mixed2']:.3e}")
        if fd is not None:
            fd["rel_l2_vs_oracle"] = chk["fd"]
            fd["cpu_baseline"] = chk["fd_cpu"]
            fd["speedup_vs_cpu_1thread"] = fd["mode_bins_per_s"] / chk["fd_cpu"]["value"]
            log(f"[bench] fd rel-L2 vs oracle ({fd_chk['modes']} modes) {chk['fd']:.3e}, "
                f"CPU {chk['fd_cpu']['value']:.4g} mode-bins/s")
        if aak is not None:
            aak["rel_l2_vs_oracle"] = chk["aak"]
            aak["cpu_baseline"] = chk["aak_cpu"]
            aak["speedup_vs_cpu_1thread"] = aak["samples_per_s"] / chk["aak_cpu"]["value"]
            amp["bitwise_vs_oracle"] = chk["interp2d"]
            amp["cpu_baseline"] = chk["interp2d_cpu"]
            amp["speedup_vs_cpu_1thread"] = amp["evaluations_per_s"] / chk["interp2d_cpu"]["value"]
            log(f"[bench] aak rel-L2 vs oracle {chk['aak']:.3e}, CPU {chk['aak_cpu']['value']:.4g} "
                f"samples/s; interp2D bitwise {chk['interp2d']}, CPU "
                f"{chk['interp2d_cpu']['value']:.4g} evaluations/s")
        cpu = cpu_baseline_batch(batch_inputs)
        log(f"[bench] CPU {cpu['value']:.4g} mode-samples/s (1 thread), "
            f"{cpu['omp']['value']:.4g} with {cpu['omp']['threads']} threads ({cpu['omp']['cpu_model']})")
        if single is not None:
            single["cpu_baseline"] = cpu_baseline(inp1)

    out = {
        "metric": "mode-samples/s & waveforms/s of the batched interpolated mode sum at rel-L2 "
                  "< 1e-6 vs the FP64 oracle (and vs the plain FP64 GPU sum)",
        "value": batch["mode_samples_per_s"],
        "unit": "mode-samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": batch["ms_per_step"],
        "timing": "steps replayed as one captured HIP graph (every waveform's splines, plans + "
                  "tiles, the persistent sum writing every sample); stage_ms from eager "
                  "launches with HIP events on the launching stream",
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 phases + split-f16 MFMA (f32 acc)",
        "data": "synthetic, seeded (fastemriwaveforms_amd/synthetic.py)",
        "config": {"workload": f"{batch['waveforms']} uneven waveforms, durations "
                               "U(0.1, 2) yr at dt = 10 s, knots U{50..200}, modes U{50..3000} "
                               f"of the 3843-mode set; {batch['samples']} samples, "
                               f"{batch['mode_samples']:.4g} mode-samples, "
                               f"{batch['intervals']} knot intervals",
                   "precision": "mixed16", "mode_samples": batch["mode_samples"],
                   "waveforms": batch["waveforms"]},
        "waveforms_per_s": batch["waveforms_per_s"],
        "rel_l2_vs_fp64_all_samples": (batch.get("fp64_control") or {}).get(
            "mixed16_rel_l2_vs_fp64_all_samples"),
        "rel_l2_vs_oracle": batch["rel_l2_vs_oracle_max_of_8"],
        "speedup_vs_fp64_plain_gpu": batch.get("speedup_vs_fp64_control"),
        "batch": batch,
        "single_waveform": single,
        "mixed2": mixed2,
        "adjoint": adjoint,
        "fd": fd,
        "aak": aak,
        "interp2d": amp,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

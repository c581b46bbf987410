#!/usr/bin/env python3
"""Benchmark of the interpolated mode sum on one MI355X.

Workload (default): full1 -- T = 1 yr, dt = 10 s (3,155,815 samples), 120
knots, all 3843 modes (l = 2..10, m = 0..l, n = -30..30), synthetic seeded
inputs resident in HBM.  One step = zero the output + spline build (2K+3
systems) + interval planning + mode sum, i.e. the whole path from knot
values to the complex128 waveform.

Prints ONE JSON line (rank 0).  value = mode-samples/s = K*N / step time.
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

PEAKS = {"fp64": 78.6, "fp32": 157.3, "hbm": 8000.0}  # TFLOP/s, TFLOP/s, GB/s (MI355X)
FLOP_PER_MODE_SAMPLE = 37  # algorithmic FLOP of the direct sum per mode-sample (sincos counted 0)


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(inp, seconds_target: float = 12.0):
    """The oracle (single thread, the reference CPU backend is single-threaded) on a
    bounded sample of the same workload: evenly spread output samples, all modes."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle
    from helpers import oracle_coefficients
    interp_tl, pcoef = oracle_coefficients(inp)
    args = (interp_tl, inp.t, pcoef, inp.m, inp.k, inp.n, inp.L, inp.num_pts, inp.K, inp.ylms,
            inp.dt, inp.t)
    # calibrate on a small sample, then size the timed sample to ~seconds_target
    probe = np.linspace(0, inp.num_pts - 1, 200).astype(np.int64)
    t0 = time.perf_counter()
    oracle.waveform_samples(probe, *args, nthreads=1)
    per = (time.perf_counter() - t0) / len(probe)
    nsamp = int(min(inp.num_pts, max(1000, seconds_target / max(per, 1e-9))))
    idx = np.linspace(0, inp.num_pts - 1, nsamp).astype(np.int64)
    t0 = time.perf_counter()
    oracle.waveform_samples(idx, *args, nthreads=1)
    el = time.perf_counter() - t0
    rate = nsamp * inp.K / el
    return {"value": rate, "unit": "mode-samples/s", "cores": 1, "kind": "port",
            "sample": f"{nsamp} evenly spread output samples x {inp.K} modes of the same workload "
                      f"({nsamp * inp.K:.3g} mode-samples, {el:.1f} s), oracle/few_oracle.c "
                      f"restating interpolate.cu's CPU branch, -O2 -ffp-contract=off"}


def rel_l2_vs_oracle(inp, h, count=1500):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle
    from helpers import oracle_waveform, spread_indices
    idx = spread_indices(inp.num_pts, count=count, seed=5)
    ref = oracle_waveform(inp, idx, nthreads=min(16, os.cpu_count() or 1))
    return oracle.rel_l2(h[idx], ref)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="full1")
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the sum shards by "
              "time with no exchange; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    from fastemriwaveforms_amd import synthetic
    from fastemriwaveforms_amd.pipeline import ModeSumProblem

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    log(f"[bench] building workload {args.workload}")
    inp = synthetic.make_workload(args.workload)
    prob = ModeSumProblem(inp, dev)
    K, N = inp.K, inp.num_pts
    log(f"[bench] K={K} N={N} L={inp.L} mode-samples={K * N:.4g}")

    for _ in range(args.warmup):
        prob.run(args.precision)
    torch.cuda.synchronize()
    log("[bench] warmup done")

    # live per-kernel timing of the dominant kernel (mode sum) with events on
    # the stream the kernels are launched on (torch's current stream)
    ev_sum = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.steps)]
    ev_spl = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        prob.wave.zero_()
        ev_spl[s][0].record()
        prob.build_splines()
        ev_spl[s][1].record()
        ev_sum[s][0].record()
        prob.mode_sum(args.precision)
        ev_sum[s][1].record()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms_step = el / args.steps * 1e3
    sum_ms = float(np.mean([a.elapsed_time(b) for a, b in ev_sum]))
    spl_ms = float(np.mean([a.elapsed_time(b) for a, b in ev_spl]))
    value = K * N / (el / args.steps)
    log(f"[bench] step {ms_step:.3f} ms, mode-sum {sum_ms:.3f} ms, spline {spl_ms:.4f} ms")

    h = prob.wave.cpu().numpy()
    err = rel_l2_vs_oracle(inp, h)
    log(f"[bench] rel-L2 vs oracle {err:.3e}")

    flops = FLOP_PER_MODE_SAMPLE * K * N
    achieved = flops / (sum_ms * 1e-3) / 1e12
    peak = PEAKS["fp64"]
    roofline = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None,
                "kernel": "direct_fp64_kernel", "pipe": "FP64 VALU (no MFMA in this kernel)",
                "algorithmic": f"{FLOP_PER_MODE_SAMPLE} FLOP/mode-sample x {K * N:.4g} "
                               "mode-samples per launch, sincos counted 0"}

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle, 1 thread)")
        cpu = cpu_baseline(inp)
        log(f"[bench] CPU {cpu['value']:.4g} mode-samples/s")

    out = {
        "metric": "mode-samples/s (interpolated mode sum, spline build included)",
        "value": value,
        "unit": "mode-samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded; knots, amplitudes, phases, Y_lm as described in synthetic.py)",
        "config": {"workload": f"{args.workload}: T={inp.T_years} yr, dt={inp.dt} s, N={N} samples, "
                               f"L={inp.L} knots, K={K} modes",
                   "precision": args.precision, "mode_samples": K * N},
        "waveforms_per_s": 1e3 / ms_step,
        "rel_l2_vs_oracle": err,
        "kernel_ms": {"mode_sum": sum_ms, "spline_build": spl_ms},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

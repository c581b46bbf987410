"""Project autotuner for the mode-sum kernels.

Times each candidate (precision, variant) of the summation stage on the
problem at hand with HIP events, keeps the fastest one that is supported,
and caches the choice in a JSON file keyed by device name and tile shape
(row blocks, k-blocks of the mode plan).  Cache path: $FEW_AMD_AUTOTUNE_CACHE,
else ~/.cache/fastemriwaveforms_amd/autotune.json.
"""

from __future__ import annotations

import json
import os

import torch

# mixed16 variants are a bit set (include/few_rocm.h): 4 = 4 waves per
# workgroup instead of 8, 8 = no raised wave priority over the MFMA phase,
# 16 = the row-streamed contraction (taken where the rows are arithmetic and
# the plan has at most 8 column steps, else the regular kernel runs), 32 =
# twice as many persistent workgroups as fit at once.  All 16 are timed.
MIXED16_VARIANTS = tuple(w | p | s | g for w in (0, 4) for p in (0, 8) for s in (0, 16)
                         for g in (0, 32))
CANDIDATES = [("mixed16", v) for v in MIXED16_VARIANTS] + [("mixed", 0), ("mixed", 4)]
BATCH_VARIANTS = MIXED16_VARIANTS


def cache_path() -> str:
    p = os.environ.get("FEW_AMD_AUTOTUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "fastemriwaveforms_amd", "autotune.json")


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(d: dict) -> None:
    p = cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        with open(p + ".tmp", "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(p + ".tmp", p)
    except OSError:
        pass  # a read-only home only costs re-tuning


def _key(kind: str, a: int, b: int) -> str:
    """single: a, b = plan rows, columns; batch: a, b = total intervals, waveforms."""
    name = torch.cuda.get_device_name(torch.cuda.current_device())
    return f"{name}|{kind}|{a}|{b}"


def _time(fn, reps: int) -> float:
    fn()  # warm-up (also compiles nothing: all kernels are prebuilt)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def tune_problem(prob, reps: int = 3, use_cache: bool = True, candidates=None):
    """Pick (precision, variant) for a single-waveform ModeSumProblem; sets prob.variant
    and returns {"precision", "variant", "ms", "timings"}."""
    key = _key("single", prob.plan.rows, prob.plan.cols)
    cache = _load() if use_cache else {}
    if use_cache and key in cache:
        best = cache[key]
        prob.variant = best["variant"]
        return best
    timings = {}
    for prec, var in candidates or CANDIDATES:
        if (prec == "mixed16" and not prob.plan.supported16) or (prec == "mixed" and not prob.plan.supported):
            continue
        prob.variant = var
        prob.wave.zero_()
        prob.build_splines()
        prob.mode_sum(prec, stage=1)
        timings[f"{prec}/{var}"] = _time(lambda: prob.mode_sum(prec, stage=2), reps)
    if not timings:
        raise RuntimeError("no supported candidate for this mode set")
    name = min(timings, key=timings.get)
    prec, var = name.split("/")
    best = {"precision": prec, "variant": int(var), "ms": timings[name], "timings": timings}
    prob.variant = int(var)
    if use_cache:
        cache[key] = best
        _store(cache)
    return best


def tune_batch(bp, reps: int = 2, use_cache: bool = True):
    """Pick the variant of the batched mixed16 sum; sets bp.variant."""
    from .pipeline import BATCH_STAGE_PLAN, BATCH_STAGE_SPLINES, BATCH_STAGE_SUM
    key = _key("batch", int(bp.totals[8]), len(bp.inputs)) + f"|ms{bp.mode_samples:.2e}"
    cache = _load() if use_cache else {}
    if use_cache and key in cache:
        bp.variant = cache[key]["variant"]
        return cache[key]
    timings = {}
    for var in BATCH_VARIANTS:
        bp.variant = var
        bp.wave.zero_()
        bp.stages("mixed16", BATCH_STAGE_SPLINES | BATCH_STAGE_PLAN)
        timings[f"mixed16/{var}"] = _time(lambda: bp.stages("mixed16", BATCH_STAGE_SUM), reps)
    name = min(timings, key=timings.get)
    best = {"precision": "mixed16", "variant": int(name.split("/")[1]), "ms": timings[name],
            "timings": timings}
    bp.variant = best["variant"]
    if use_cache:
        cache[key] = best
        _store(cache)
    return best

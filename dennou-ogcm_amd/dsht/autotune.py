"""Autotuner for the per-plan launch knobs (dsht_set_tuning).

For one plan and a representative column count it times the dynamics-step transform set
(value and both gradients of 4 fields in one fused inverse call, 1 forward over 6
tendencies) under each candidate and keeps the fastest.  The candidates are:
  column chunk      TUNE_CHUNK_COLS: columns per Fourier scratch chunk (0 = whole call up to
                    512 MiB of Fourier scratch per output)
  column block      TUNE_LEG_CB: column tiles per wave / workgroup of the contractions
  contraction       TUNE_LEG: kernel family (split-form inverse: per form, packed data
                    panels or whole-K LDS staging; FP64: LDS-staged or per-wave fragment
                    streams)
Results are cached per (nmax, im, jm, precision, ncols bucket) in a JSON file so later
runs reuse them.
"""
import json
import os
import time

from . import PREC_FP64, TUNE_CHUNK_COLS, TUNE_LEG, TUNE_LEG_CB, VAL

CACHE = os.environ.get("DSHT_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "dsht", "tuning.json"))
KNOBS = (("chunk", TUNE_CHUNK_COLS), ("cb", TUNE_LEG_CB), ("leg", TUNE_LEG))


def _bucket(ncols):
    b = 16
    while b < ncols:
        b *= 2
    return b


def _key(plan, ncols):
    return f"{plan.nmax}x{plan.im}x{plan.jm}:p{plan.precision}:c{_bucket(ncols)}"


def candidates(plan, ncols):
    """Candidate knob settings as dicts {chunk, cb, leg}."""
    chunks = sorted({c for c in (64, 128, 256) if c < ncols} | {0})
    legs = (0, 1, 2) if plan.precision == PREC_FP64 else (0, 2, 3)
    return [{"chunk": c, "cb": cb, "leg": lg} for c in chunks for cb in (16, 32) for lg in legs]


def apply(plan, setting):
    for name, key in KNOBS:
        plan.set_tuning(key, int(setting.get(name, 0)))


def _time_step(plan, bufs, stream, reps=3):
    import torch
    spec, grids, tend, spect = bufs
    for _ in range(2):
        plan.inv_vgrad_d(spec, *grids, stream=stream)
        plan.fwd_d(VAL, tend, spect, stream=stream)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        plan.inv_vgrad_d(spec, *grids, stream=stream)
        plan.fwd_d(VAL, tend, spect, stream=stream)
    torch.cuda.synchronize()
    return (time.perf_counter() - t) / reps


def tune(plan, ncols, dev=None, reps=3, use_cache=True, verbose=False):
    """Pick the knob setting for `plan` at `ncols` prognostic columns; applies and returns
    {chunk, cb, leg, ms, candidates}."""
    import torch

    cache = {}
    if use_cache and os.path.exists(CACHE):
        try:
            cache = json.load(open(CACHE))
        except (OSError, ValueError):
            cache = {}
    key = _key(plan, ncols)
    if key in cache:
        apply(plan, cache[key])
        return cache[key]
    dev = dev or torch.device("cuda:0")
    f64 = dict(dtype=torch.float64, device=dev)
    spec = torch.randn((plan.lm, ncols), **f64) * 1e-3
    grids = [torch.empty((ncols, plan.jm, plan.im), **f64) for _ in range(3)]
    nt = max(1, ncols * 3 // 2)
    tend = torch.randn((nt, plan.jm, plan.im), **f64)
    spect = torch.empty((plan.lm, nt), **f64)
    stream = torch.cuda.current_stream(dev).cuda_stream
    best = None
    results = []
    for cand in candidates(plan, ncols):
        apply(plan, cand)
        dt = _time_step(plan, (spec, grids, tend, spect), stream, reps)
        results.append(dict(cand, ms=1e3 * dt))
        if verbose:
            print(f"  tune {key}: {cand} {1e3 * dt:.3f} ms", flush=True)
        if best is None or dt < best[0]:
            best = (dt, cand)
    apply(plan, best[1])
    entry = dict(best[1], ms=1e3 * best[0], candidates=results)
    if use_cache:
        cache[key] = entry
        try:
            os.makedirs(os.path.dirname(CACHE), exist_ok=True)
            json.dump(cache, open(CACHE, "w"), indent=1)
        except OSError:
            pass
    return entry
